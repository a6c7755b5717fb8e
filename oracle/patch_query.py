"""Print TOPAZ's query.c with its two per-hit loops replaced by one call each
to the batch entry points of libtopaz_amd (include/topaz_ssw_compat.h).

    python3 oracle/patch_query.py /root/reference/src/query.c | gcc -x c -c - ...

This is the source change a maintainer would make to use the batch ABI with
TOPAZ's own types.  The patched text goes to stdout only: oracle/Makefile
compiles it straight into oracle/_ref/batch/query_batch.o, which the drop-in
test links in place of query.o.

The two loops are found by the calls they make, not by line numbers:
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
"""

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
    a0, a1 = function_body(lines, net, "query_perform_alignments")
    s0, s1 = function_body(lines, net, "query_perform_statistics")
    la = loop_with_call(lines, net, a0, a1, "while(", "perform_alignment_with_profile")
    ls = loop_with_call(lines, net, s0, s1, "for(", "stats_using_ssw_profile")
    assert a0 < la[0] < la[1] <= a1 <= s0 < ls[0] < ls[1] <= s1
    # the declarations go just before query_perform_alignments
    out = lines[:a0] + DECLS.split("\n") + lines[a0:la[0]] + ALIGN.rstrip("\n").split("\n") + \
        lines[la[1]:ls[0]] + STATS.rstrip("\n").split("\n") + lines[ls[1]:]
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1])
