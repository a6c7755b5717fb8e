"""Print TOPAZ's suffix.c with the seed walk's gapless extensions prefetched
in batches through libtopaz_amd's sa_hsp_batch.

    python3 oracle/patch_suffix.py /root/reference/src/suffix.c | gcc -x c -c - ...

A suffix walks the suffix array one slot at a time (suffix_increment,
suffix.c:38-52, driven by pqueue_next, priority_queue.c:37-56), extending
each slot's seed with one sa_hsp call.  The extension of slot p depends only
on (query offset, p), so the walk can compute the next slots' extensions
ahead of time without changing anything it does: the patched suffix keeps a
cache of the extensions of the next TPZ_WALK_K slots in its direction,
filled by one sa_hsp_batch call when it runs out (or up front by the patched
query_find_suffixes, patch_query.py).  The cache sits behind the suffix_t in
the same allocation, so suffix.h and the other objects are unchanged.

suffix_alloc, suffix_duplicate, suffix_free and suffix_increment (found by
their definitions, not by line numbers) are replaced by the code below, and
a helper, suffix_prefetched, is added for query_find_suffixes.  A suffix's
refills grow from TPZ_WALK_K slots, doubling each time.  The patched
text goes to stdout only: oracle/Makefile compiles it straight into
oracle/_ref/batch/suffix_batch.o.
"""
import sys

HEADER = """
/* ---- seed-walk prefetch through libtopaz_amd (include/topaz_ssw_compat.h) ---- */
void sa_hsp_batch(db_t* db, seq_t* q, int n, const int* qoffsets, const int64_t* positions,
                  options_t* opt, int* scores, int* lengths);

/* a refill extends TPZ_WALK_K slots, twice as many as the suffix's last
   refill after that, up to TPZ_WALK_MAX: a walk that has gone far tends to
   go on, and each refill is a GPU round trip */
#define TPZ_WALK_K 128
#define TPZ_WALK_MAX 16384

/* extensions of the slots first, first + direction, ..., n of them */
typedef struct {
    saidx64_t first;
    int n, next, cap, want;
    int *score, *length;
} tpz_walk_cache;

static tpz_walk_cache* tpz_cache(suffix_t* suf) {
    return (tpz_walk_cache*) (suf + 1);
}

static suffix_t* tpz_suffix_new(void) {
    suffix_t* suf = (suffix_t*) malloc(sizeof(suffix_t) + sizeof(tpz_walk_cache));
    tpz_walk_cache* c = tpz_cache(suf);
    c->n = 0;
    c->next = 0;
    c->cap = 0;
    c->want = TPZ_WALK_K;
    c->score = NULL;
    c->length = NULL;
    return suf;
}

static void tpz_reserve(tpz_walk_cache* c, int n) {
    if(n > c->cap) {
        c->score = (int*) realloc(c->score, sizeof(int) * n);
        c->length = (int*) realloc(c->length, sizeof(int) * n);
        c->cap = n;
    }
}

/* the extensions of the n slots after suf->position in its direction,
   computed by the caller (score as sa_hsp writes it: only where length != -1) */
void suffix_prefetched(suffix_t* suf, int n, const int* scores, const int* lengths) {
    tpz_walk_cache* c = tpz_cache(suf);
    int k;
    tpz_reserve(c, n);
    c->first = suf->position + suf->direction;
    c->n = n;
    c->next = 0;
    for(k = 0; k < n; ++k) {
        c->score[k] = scores[k];
        c->length[k] = lengths[k];
    }
}
"""

ALLOC = """suffix_t* suffix_alloc(seq_t* seq, db_t *db, int offset, saidx64_t position, int direction, int gapless_score, int gapless_length) {
    suffix_t* suf = tpz_suffix_new();
    (void) db;
    suf->seq = seq;
    suf->offset = offset;
    suf->position = position;
    suf->direction = direction;
    suf->gapless_score = gapless_score;
    suf->gapless_length = gapless_length;
    return suf;
}"""

DUPLICATE = """suffix_t* suffix_duplicate(suffix_t* suf) {
    suffix_t* tmp = tpz_suffix_new();
    memcpy(tmp, suf, sizeof(suffix_t));
    tmp->direction *= -1;
    return tmp;
}"""

FREE = """void suffix_free(suffix_t *suf) {
    if(suf) {
        free(tpz_cache(suf)->score);
        free(tpz_cache(suf)->length);
        free(suf);
    }
}"""

INCREMENT = """int suffix_increment(suffix_t *suf, db_t* db, options_t* opt) {
    tpz_walk_cache* c = tpz_cache(suf);
    int j;

    suf->position += suf->direction;

    if(!db_valid_index(db, suf->position))
        return -1;

    if(suf->position >= (saidx64_t) db->Tsize) {
        /* past the text: no batch; the single call behaves as before */
        sa_hsp(db, suf->seq, suf->offset, suf->position, opt,
               &suf->gapless_score, &suf->gapless_length);
        return 0;
    }

    if(c->next >= c->n || c->first + (saidx64_t) c->next * suf->direction != suf->position) {
        /* refill: this slot and the next ones in the walk's direction */
        const int want = c->want;
        int* offs = (int*) malloc(sizeof(int) * want);
        int64_t* pos = (int64_t*) malloc(sizeof(int64_t) * want);
        saidx64_t p = suf->position;
        int k = 0;
        while(k < want && p >= 0 && p < (saidx64_t) db->Tsize) {
            offs[k] = suf->offset;
            pos[k] = p;
            ++k;
            p += suf->direction;
        }
        tpz_reserve(c, k);
        sa_hsp_batch(db, suf->seq, k, offs, pos, opt, c->score, c->length);
        free(offs);
        free(pos);
        c->first = suf->position;
        c->n = k;
        c->next = 0;
        c->want = want < TPZ_WALK_MAX ? 2 * want : TPZ_WALK_MAX;
    }

    j = c->next++;
    suf->gapless_length = c->length[j];
    if(c->length[j] != -1)            /* sa_hsp leaves the score alone then (search.c:113-116) */
        suf->gapless_score = c->score[j];
    return 0;
}"""


def live_braces(lines):
    return [line.count("{") - line.count("}") if not line.strip().startswith("#") else 0
            for line in lines]


def definition(lines, net, head):
    """[start, end) of the function whose definition line starts with head."""
    for k, line in enumerate(lines):
        if line.startswith(head):
            depth = 0
            for j in range(k, len(lines)):
                depth += net[j]
                if depth == 0:
                    return k, j + 1
    raise SystemExit("patch_suffix: no definition starting with %r" % head)


def main(path):
    lines = open(path).read().split("\n")
    net = live_braces(lines)
    spans = [(definition(lines, net, head), body) for head, body in (
        ("suffix_t* suffix_alloc(", ALLOC),
        ("suffix_t* suffix_duplicate(", DUPLICATE),
        ("void suffix_free(", FREE),
        ("int suffix_increment(", INCREMENT))]
    spans.sort()
    out, at = [], 0
    for (lo, hi), body in spans:
        out += lines[at:lo] + body.split("\n")
        at = hi
    out += lines[at:]
    last_inc = max(k for k, line in enumerate(out) if line.startswith("#include"))
    out = out[:last_inc + 1] + HEADER.split("\n") + out[last_inc + 1:]
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1])
