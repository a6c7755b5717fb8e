/*
 * cpu_bench -- TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py, the
 * reference's own compiled ssw.c and search.c timed on all host cores.
 *
 *   cpu_bench sw      FILE THREADS SECONDS
 *   cpu_bench gapless FILE THREADS SECONDS
 *
 * sw: FILE holds queries, subjects and pairs (grouped by query); each
 * thread takes the next query, builds its profile once (ssw_init with
 * score_size 2, as get_ssw_profile does, ssw_wrapper.c:105-130, query.c:194-197)
 * and aligns all its pairs with ssw_align as align_using_ssw_profile calls it
 * (ssw_wrapper.c:209-218), until SECONDS have passed.  Prints
 *   cells pairs queries seconds
 *
 * gapless: FILE holds a text, queries and seeds (query, offset, text
 * position); threads take chunks of 65,536 seeds in order, going round the
 * list again at its end, and run the reference's sa_hsp (search.c:99-190)
 * on each, until SECONDS have passed; the chunks taken are complete, so the
 * seeds done are seeds 0 .. done-1 taken modulo the list.  Prints
 *   seeds seconds
 *
 * File layouts (little endian) are written by bench.py.  Built by
 * oracle/Makefile against the reference's objects; never shipped.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "db.h"
#include "options.h"
#include "search.h"
#include "seq.h"
#include "ssw.h"
#include "substitution.h"

static double now(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void* slurp(const char* path, size_t* size) {
    FILE* f = fopen(path, "rb");
    char* buf;
    if (!f) { perror(path); exit(1); }
    fseek(f, 0, SEEK_END);
    *size = (size_t) ftell(f);
    fseek(f, 0, SEEK_SET);
    buf = (char*) malloc(*size + 1);
    if (fread(buf, 1, *size, f) != *size) { perror(path); exit(1); }
    fclose(f);
    return buf;
}

/* ---- sw ------------------------------------------------------------------ */
typedef struct {
    int64_t n_q, n_s, n_pairs, q_total, s_total;
    const int64_t *q_off, *s_off;
    const int32_t *q_len, *s_len, *pq, *ps;
    const int8_t *q_res, *s_res;
    const int64_t* group;          /* first pair of each query group, n_groups + 1 */
    int64_t n_groups;
    double deadline;
    pthread_mutex_t mu;
    int64_t next, cells, pairs, queries;
} sw_job;

static void* sw_worker(void* arg) {
    sw_job* j = (sw_job*) arg;
    for (;;) {
        int64_t g, p, cells = 0;
        s_profile* prof;
        int32_t qi;
        pthread_mutex_lock(&j->mu);
        g = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (g >= j->n_groups || now() > j->deadline) break;
        qi = j->pq[j->group[g]];
        prof = ssw_init(j->q_res + j->q_off[qi], j->q_len[qi], BLOSUM62, 25, 2);
        for (p = j->group[g]; p < j->group[g + 1]; ++p) {
            const int32_t si = j->ps[p];
            s_align* a = ssw_align(prof, j->s_res + j->s_off[si], j->s_len[si], 11, 1, 0, 0, 0, 15);
            if (a) align_destroy(a);
            cells += (int64_t) j->q_len[qi] * j->s_len[si];
        }
        init_destroy(prof);
        pthread_mutex_lock(&j->mu);
        j->cells += cells;
        j->pairs += j->group[g + 1] - j->group[g];
        j->queries += 1;
        pthread_mutex_unlock(&j->mu);
    }
    return NULL;
}

static int run_sw(const char* path, int threads, double seconds) {
    size_t size;
    char* b = (char*) slurp(path, &size);
    const int64_t* h = (const int64_t*) b;
    sw_job j;
    char* c;
    int64_t p, n;
    pthread_t* th;
    int t;
    double t0;
    memset(&j, 0, sizeof j);
    j.n_q = h[0]; j.n_s = h[1]; j.n_pairs = h[2]; j.q_total = h[3]; j.s_total = h[4];
    c = b + 5 * 8;
    j.q_off = (const int64_t*) c; c += 8 * j.n_q;
    j.s_off = (const int64_t*) c; c += 8 * j.n_s;
    j.q_len = (const int32_t*) c; c += 4 * j.n_q;
    j.s_len = (const int32_t*) c; c += 4 * j.n_s;
    j.pq = (const int32_t*) c; c += 4 * j.n_pairs;
    j.ps = (const int32_t*) c; c += 4 * j.n_pairs;
    j.q_res = (const int8_t*) c; c += j.q_total;
    j.s_res = (const int8_t*) c;
    {
        int64_t* group = (int64_t*) malloc(8 * (j.n_pairs + 2));
        n = 0;
        for (p = 0; p < j.n_pairs; ++p)
            if (p == 0 || j.pq[p] != j.pq[p - 1]) group[n++] = p;
        group[n] = j.n_pairs;
        j.group = group;
        j.n_groups = n;
    }
    pthread_mutex_init(&j.mu, NULL);
    th = (pthread_t*) malloc(sizeof(pthread_t) * threads);
    t0 = now();
    j.deadline = t0 + seconds;
    for (t = 0; t < threads; ++t) pthread_create(&th[t], NULL, sw_worker, &j);
    for (t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    printf("%lld %lld %lld %.6f\n", (long long) j.cells, (long long) j.pairs,
           (long long) j.queries, now() - t0);
    return 0;
}

/* ---- gapless ------------------------------------------------------------- */
typedef struct {
    db_t db;
    options_t opt;
    seq_t** queries;
    const int32_t *sq, *so;
    int64_t n_seeds;
    double deadline;
    pthread_mutex_t mu;
    int64_t next, done;
} g_job;

#define G_CHUNK 65536

static void* g_worker(void* arg) {
    g_job* j = (g_job*) arg;
    for (;;) {
        int64_t k0, k;
        int score, length;
        pthread_mutex_lock(&j->mu);
        k0 = j->next;
        if (now() <= j->deadline) j->next += G_CHUNK;
        else k0 = -1;
        pthread_mutex_unlock(&j->mu);
        if (k0 < 0) break;
        for (k = k0; k < k0 + G_CHUNK; ++k) {
            /* seed i's text position is db.SA[i] */
            const int64_t i = k % j->n_seeds;
            sa_hsp(&j->db, j->queries[j->sq[i]], j->so[i], i, &j->opt, &score, &length);
        }
        pthread_mutex_lock(&j->mu);
        j->done += G_CHUNK;
        pthread_mutex_unlock(&j->mu);
    }
    return NULL;
}

static int run_gapless(const char* path, int threads, double seconds) {
    size_t size;
    char* b = (char*) slurp(path, &size);
    const int64_t* h = (const int64_t*) b;
    const int64_t t_size = h[0], n_q = h[1], q_total = h[2], n_seeds = h[3];
    const int64_t *q_off, *pos;
    const int32_t* q_len;
    const char* qc;
    char* c = b + 4 * 8;
    g_job j;
    int64_t k;
    pthread_t* th;
    int t;
    double t0;
    memset(&j, 0, sizeof j);
    j.db.T = c; c += t_size;
    j.db.Tsize = (uint64_t) t_size;
    q_off = (const int64_t*) c; c += 8 * n_q;
    q_len = (const int32_t*) c; c += 4 * n_q;
    qc = c; c += q_total;
    j.sq = (const int32_t*) c; c += 4 * n_seeds;
    j.so = (const int32_t*) c; c += 4 * n_seeds;
    pos = (const int64_t*) c;
    j.db.SA = (saidx64_t*) pos;                  /* slot k holds seed k's position */
    j.db.SAsize = (uint64_t) n_seeds * 8;
    j.n_seeds = n_seeds;
    j.opt.substitution = &BLOSUM62;
    j.opt.min_hit_length = DEFAULT_MINLENGTH;
    j.queries = (seq_t**) malloc(sizeof(seq_t*) * n_q);
    for (k = 0; k < n_q; ++k) {
        j.queries[k] = seq_alloc();
        seq_seq(j.queries[k], (char*) qc + q_off[k], q_len[k], 0);
    }
    pthread_mutex_init(&j.mu, NULL);
    th = (pthread_t*) malloc(sizeof(pthread_t) * threads);
    t0 = now();
    j.deadline = t0 + seconds;
    for (t = 0; t < threads; ++t) pthread_create(&th[t], NULL, g_worker, &j);
    for (t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    printf("%lld %.6f\n", (long long) j.done, now() - t0);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: cpu_bench sw|gapless FILE THREADS SECONDS\n");
        return 2;
    }
    if (strcmp(argv[1], "sw") == 0) return run_sw(argv[2], atoi(argv[3]), atof(argv[4]));
    if (strcmp(argv[1], "gapless") == 0) return run_gapless(argv[2], atoi(argv[3]), atof(argv[4]));
    return 2;
}
