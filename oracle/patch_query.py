"""Print TOPAZ's query.c with its per-seed and per-hit loops replaced by
batch calls to libtopaz_amd (include/topaz_ssw_compat.h).

    python3 oracle/patch_query.py /root/reference/src/query.c | gcc -x c -c - ...

This is the source change a maintainer would make to use the batch ABI with
TOPAZ's own types.  The patched text goes to stdout only: oracle/Makefile
compiles it straight into oracle/_ref/batch/query_batch.o, which the drop-in
test links in place of query.o.

The code is found by definitions and calls, not by line numbers:
  * query_find_suffixes (query.c:94-131) is rewritten: one sa_search_batch
    call for the suffix-array search of every offset, then one sa_hsp_batch
    call for both
    seeds of every offset and the first slots of each new suffix's walk
    (FIND_K, more in a larger text; handed to the suffix with
    suffix_prefetched, patch_suffix.py),
    then the suffixes in the original order;
  * in query_perform_alignments (query.c:178-248), the `while` loop over the
    hit-list traversal that calls perform_alignment_with_profile per hit
    (query.c:202-228);
  * in query_perform_statistics (query.c:250-299), the `for` loop over the
    top hits that calls stats_using_ssw_profile per hit (query.c:276-291).
Each is replaced, from its first line to its matching closing brace, by the
code below.
"""
import sys

DECLS = """
/* batch entry points of libtopaz_amd (include/topaz_ssw_compat.h) */
int align_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                  options_t* opt);
void stats_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                   options_t* opt);
void sa_hsp_batch(db_t* db, seq_t* q, int n, const int* qoffsets, const int64_t* positions,
                  options_t* opt, int* scores, int* lengths);
int sa_search_batch(db_t* db, seq_t* q, int n, const int* offsets, int64_t* idx,
                    int64_t* num_matches, int64_t* last_position);
/* suffix_batch.o (patch_suffix.py): the next extensions of a new suffix's walk */
void suffix_prefetched(suffix_t* suf, int n, const int* scores, const int* lengths);
"""

# slots of each new suffix's walk prefetched with the first extensions: at
# least FIND_K, doubled while FIND_K slots per 1.4 M text bytes is more
# (walks run longer in a larger database), at most FIND_K_MAX
FIND_K = 64
FIND_K_MAX = 1024

FIND = """int query_find_suffixes(query_t* q) {
    /* 1. the suffix-array interval of every query offset, in one batch */
    int i, h, j, n_cand = 0, n_hit = 0, n_seed = 0;
    int num_suffixes = seq_len(q->query) - q->opt->min_hit_length;
    const saidx64_t tsize = (saidx64_t) q->db->Tsize;
    int K = %(K)d;
    size_t cap = num_suffixes > 0 ? (size_t) num_suffixes : 1;
    int* hit_off = malloc(sizeof(int) * cap);
    saidx64_t* hit_pos = malloc(sizeof(saidx64_t) * cap);
    int *offs, *scores, *lengths, *first;
    int64_t* slots;

    while(K < %(KMAX)d && (saidx64_t) K * 1400000 < tsize)
        K *= 2;
    for(i = 0; i < num_suffixes; ++i) {
        if(_bad_prefix(seq_ptr(q->query) + i))
            continue;
        hit_off[n_cand++] = i;
    }
    if(n_cand > 0 && sa_search_batch(q->db, q->query, n_cand, hit_off, hit_pos, NULL, NULL) != 0) {
        fprintf(stderr, "Error: internal error in suffix array search\\n");
        _exit(EXIT_FAILURE);
    }
    for(i = 0; i < n_cand; ++i) {
        if(hit_pos[i] != -1) {
            hit_off[n_hit] = hit_off[i];
            hit_pos[n_hit] = hit_pos[i];
            ++n_hit;
        }
    }

    /* 2. one batch: per interval, the forward suffix's slot pos and the K
       after it, and the backward suffix's slot pos - 1 and the K before it */
    cap = (size_t) n_hit * 2 * (K + 1) + 1;
    offs = malloc(sizeof(int) * cap);
    slots = malloc(sizeof(int64_t) * cap);
    scores = calloc(cap, sizeof(int));
    lengths = malloc(sizeof(int) * cap);
    first = malloc(sizeof(int) * (2 * (size_t) n_hit + 1));
    for(h = 0; h < n_hit; ++h) {
        first[2 * h] = n_seed;
        for(j = 0; j <= K && hit_pos[h] + j < tsize; ++j) {
            offs[n_seed] = hit_off[h];
            slots[n_seed++] = hit_pos[h] + j;
        }
        first[2 * h + 1] = n_seed;
        for(j = 0; j <= K && hit_pos[h] != 0 && hit_pos[h] - 1 - j >= 0; ++j) {
            offs[n_seed] = hit_off[h];
            slots[n_seed++] = hit_pos[h] - 1 - j;
        }
    }
    first[2 * n_hit] = n_seed;
    if(n_seed > 0)
        sa_hsp_batch(q->db, q->query, n_seed, offs, slots, q->opt, scores, lengths);

    /* 3. the suffixes, in the order the per-offset loop created them */
    for(h = 0; h < n_hit; ++h) {
        int d;
        for(d = 0; d < 2; ++d) {
            const int k = first[2 * h + d], n = first[2 * h + d + 1] - k;
            suffix_t* suf;
            if(n == 0 || lengths[k] < q->opt->min_hit_length)
                continue;
            suf = suffix_alloc(q->query, q->db, hit_off[h], slots[k], d ? -1 : 1,
                               scores[k], lengths[k]);
            suffix_prefetched(suf, n - 1, scores + k + 1, lengths + k + 1);
            pqueue_add(q->pq, suf);
        }
    }

    free(hit_off); free(hit_pos); free(offs); free(slots); free(scores); free(lengths);
    free(first);
    q->state = QUERY_GET_ALIGNMENTS;
    return 0;
}""" % {"K": FIND_K, "KMAX": FIND_K_MAX}

ALIGN = """    {
        /* every hit not yet aligned, in traversal order, ranked as before */
        size_t n_todo = 0, cap = hitlist_len(q->hl);
        hit_t** todo = malloc(sizeof(hit_t*) * (cap ? cap : 1));
        while((h = (hit_t*) traversal_next(tr)) != NULL) {
            if(!h->is_aligned) {
                h->alignment_rank = q->alignment_counter++;
                todo[n_todo++] = h;
            }
        }
        q->hit_counter += align_batch_using_ssw_profile(q->ssw_profile, q->db, todo,
                                                        (int) n_todo, q->opt);
        free(todo);
    }
"""

STATS = """    {
        /* the top hits within the E-value threshold, in bitscore order */
        size_t n_todo = 0;
        hit_t** todo = malloc(sizeof(hit_t*) * (aligned_hits ? aligned_hits : 1));
        for(i = 0; i < aligned_hits; ++i) {
            if(all_hits[i]->evalue <= q->opt->evalue) {
                todo[n_todo++] = all_hits[i];
            }
        }
        stats_batch_using_ssw_profile(q->ssw_profile, q->db, todo, (int) n_todo, q->opt);
        free(todo);
    }
"""


def live_braces(lines):
    """Net braces per line as compiled: the USE_PARASAIL branches (never
    defined in TOPAZ's build, query.c:20-23) count for nothing."""
    net, skip = [], False
    for line in lines:
        t = line.strip()
        if t.startswith("#ifdef USE_PARASAIL"):
            skip = True
        elif skip and (t.startswith("#else") or t.startswith("#endif")):
            skip = False
        net.append(0 if skip or t.startswith("#") else line.count("{") - line.count("}"))
    return net


def function_body(lines, net, name):
    """Line range [start, end) of the definition of function `name`."""
    for k, line in enumerate(lines):
        if line.startswith("int " + name + "(") and line.rstrip().endswith("{"):
            depth = 0
            for j in range(k, len(lines)):
                depth += net[j]
                if depth == 0:
                    return k, j + 1
    raise SystemExit("patch_query: no definition of %s" % name)


def loop_with_call(lines, net, lo, hi, head, call):
    """[start, end) of the loop starting with `head` in lines[lo:hi] that
    contains `call`, up to its matching closing brace."""
    for k in range(lo, hi):
        if lines[k].lstrip().startswith(head):
            depth, seen = 0, False
            for j in range(k, hi):
                depth += net[j]
                seen = seen or call in lines[j]
                if depth == 0 and j > k:
                    if seen:
                        return k, j + 1
                    break
    raise SystemExit("patch_query: no %s loop calling %s" % (head, call))


def main(path):
    lines = open(path).read().split("\n")
    net = live_braces(lines)
    f0, f1 = function_body(lines, net, "query_find_suffixes")
    a0, a1 = function_body(lines, net, "query_perform_alignments")
    s0, s1 = function_body(lines, net, "query_perform_statistics")
    la = loop_with_call(lines, net, a0, a1, "while(", "perform_alignment_with_profile")
    ls = loop_with_call(lines, net, s0, s1, "for(", "stats_using_ssw_profile")
    assert f1 <= a0 < la[0] < la[1] <= a1 <= s0 < ls[0] < ls[1] <= s1
    # the declarations go just before query_find_suffixes
    out = lines[:f0] + DECLS.split("\n") + FIND.split("\n") + lines[f1:la[0]] + \
        ALIGN.rstrip("\n").split("\n") + lines[la[1]:ls[0]] + \
        STATS.rstrip("\n").split("\n") + lines[ls[1]:]
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1])
