/*
 * ref_walk -- TEST INFRASTRUCTURE ONLY: the reference's own seed walk, run
 * through its compiled query.c / priority_queue.c / suffix.c / hit.c, for the
 * tests that check libtopaz_amd's walk (tpz_walk_query) and search.
 *
 *   ref_walk PREFIX QUERIES.fa SEEDS ALIGNMENTS [noseg]
 *
 * For every query, as read_data (engine.c:310-338) prepares it -- SEG, then
 * seq_2internal -- then query_alloc, query_find_suffixes and
 * query_fetch_alignment_batch (query.c:35-174), it prints
 *   Q <id> <length>
 *   S <the query's internal-alphabet letters>
 *   H <db_index> <gapless_score>      one line per hit, in hitlist_traversal order
 *   C <hits counted by the walk>
 * The options are TOPAZ's defaults (options.h), with -S SEEDS and -A ALIGNMENTS.
 * Built by oracle/Makefile against the reference's objects; never shipped.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "db.h"
#include "fasta.h"
#include "hit.h"
#include "options.h"
#include "query.h"
#include "rbtree.h"
#include "seg_wrapper.h"
#include "seq.h"
#include "substitution.h"

int main(int argc, char** argv) {
    options_t opt;
    db_t* db;
    fasta_t* f;
    seq_t* seq;

    if (argc < 5) {
        fprintf(stderr, "usage: ref_walk PREFIX QUERIES.fa SEEDS ALIGNMENTS [noseg]\n");
        return 2;
    }
    memset(&opt, 0, sizeof opt);
    opt.num_threads = 1;
    opt.num_hits = DEFAULT_HITS;
    opt.evalue = DEFAULT_EVALUE;
    opt.gap_open = DEFAULT_GAPOPEN;
    opt.gap_extend = DEFAULT_GAPEXTEND;
    opt.lambda = DEFAULT_LAMBDA;
    opt.k = DEFAULT_K;
    opt.H = DEFAULT_H;
    opt.substitution = &BLOSUM62;
    opt.min_hit_length = DEFAULT_MINLENGTH;
    opt.min_hit_score = DEFAULT_MINSCORE;
    opt.number_of_seeds = atoi(argv[3]);
    opt.number_of_alignments = atoi(argv[4]);
    opt.superfast = 1;
    opt.seg_enabled = !(argc > 5 && strcmp(argv[5], "noseg") == 0);

    db = db_alloc(argv[1]);
    if (db_load(db) < 0) {
        fprintf(stderr, "ref_walk: cannot load %s\n", argv[1]);
        return 1;
    }
    opt.db_size = db->Tsize - db->Psize;        /* engine.c:52-55 */
    opt.db_proteins = db->Psize;
    if (opt.seg_enabled)
        initialise_seg();
    f = fasta_alloc(argv[2]);
    while ((seq = fasta_next(f, NULL)) != NULL) {
        query_t* q;
        traversal_t* tr;
        hit_t* h;
        int counted = 0;

        if (opt.seg_enabled)
            run_seg(seq);
        seq_2internal(seq);
        printf("Q %.*s %d\n", seq_idlen(seq), seq_idptr(seq), seq_len(seq));
        printf("S %.*s\n", seq_len(seq), seq_ptr(seq));
        q = query_alloc(seq, db, &opt);
        query_find_suffixes(q);
        query_fetch_alignment_batch(q);
        tr = hitlist_traversal(q->hl);
        traversal_start(tr);
        while ((h = (hit_t*) traversal_next(tr)) != NULL) {
            printf("H %lu %d\n", (unsigned long) h->db_index, h->gapless_score);
            ++counted;
        }
        traversal_free(tr);
        printf("C %d\n", counted);
        query_free(q);
    }
    fasta_free(f);
    db_free(db);
    return 0;
}
