// The seed walk of one query and its hit list over fully computed streams,
// as one function compiled for both the host and the GPU: TOPAZ's priority
// queue over suffix-array streams (priority_queue.c:36-56, suffix.c:40-65)
// feeding hitlist_add (hit.c:126-193), reproduced exactly.
//
// A stream is one suffix of the reference's walk: a query offset, a first
// suffix-array slot and a direction; element k is the gapless score at slot
// first + k * direction, and the stream holds the elements before the first
// one whose extension is not a seed (length < min_hit_length) or that falls
// outside the suffix array: the reference stops the stream there.  Each
// element also carries its protein (db_index of its text position).  A
// stream may be cut short (open): a run that reaches its end stops the walk
// with kWalkNeed, and the caller computes more and walks again.
//
// Runs.  After a stream's element x is popped (x is then the largest score
// in the queue), its next element y >= x is at once the largest again, and
// the newest among equals (rbtree.c:72-91, 410-425), so it is popped next;
// the same holds for every following element >= x.  The reference therefore
// pops a stream in runs: a run starts at an element smaller than every
// earlier element of its stream (a strict prefix minimum, the run's leader)
// and takes every following element >= the leader.  Runs of one stream have
// strictly decreasing leaders.  Popping whole runs from a queue ordered by
// (leader, insertion number) -- a run inserted when the run before it in its
// stream ends -- gives the reference's pop sequence element for element, with
// one queue operation per run instead of per element.
//
// Run queue.  A run pushed while popping a run with leader L has a leader
// below L, so the leaders popped never increase; and among equal leaders the
// newest run pops first.  The queue is therefore a stack per leader score
// (top[score]) with a pointer that only moves down, and since a stream has
// at most one run waiting, the stacks are threaded through one node per
// stream (its run's first element and the stream below it on its stack).
//
// Hit list.  At most `capacity` proteins with their best gapless score, in
// the priority tree's order: (score, insertion number), a re-put counting as
// a new insertion (rbtree.c:496-505).  Per score a doubly linked list, oldest
// first; proteins found through an open-addressing table (linear probing,
// backward-shift deletion).  The walk stops once `limit` elements have been
// counted (query.c:144-150): a new protein, or the list full and its worst
// score >= the element's.
//
// All state lives in caller-provided arrays (WalkMem), laid out by
// walk_mem_layout; the head/tail and table-value arrays start as -1.  The
// GPU runs one query per wavefront with every lane executing the same
// steps on the same data (uniform loads and stores), except the scan of a
// run's elements, 64 at a time.
#ifndef TOPAZ_AMD_WALK_CORE_H
#define TOPAZ_AMD_WALK_CORE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// Stream element value of a slot that is not a seed.
constexpr int16_t kStreamEnd = -1;

#define TPZ_HD __host__ __device__ __forceinline__

struct WalkNode {
    int32_t start;               // first element of the stream's waiting run
    int32_t next;                // the stream below it on its stack, -1: none
};

struct WalkEntry {
    uint32_t db_index;
    int32_t score;
    int32_t prev, next;          // list of its score, oldest first
};

// One query's streams (element arrays in an arena shared by the queries).
struct WalkStreams {
    int32_t n;
    const int64_t* base;         // per stream: index of its element 0 in vals / prot
    const int32_t* len;          // per stream: its seeds (0: never created, query.c:119,127)
    const uint8_t* open;         // per stream: cut short, more seeds may follow (or null)
    const int16_t* vals;
    const uint32_t* prot;
};

struct WalkMem {
    int32_t* top;                // max_score + 1, -1 initially: stream on top per leader
    WalkNode* node;              // n_streams entries
    WalkEntry* entry;            // capacity + 1
    int32_t* free_stack;         // capacity + 1
    int32_t* head;               // max_score + 1, -1 initially
    int32_t* tail;               // max_score + 1, -1 initially
    uint32_t* key;               // table_size
    int32_t* val;                // table_size, -1 initially
    uint32_t table_mask;         // table_size - 1 (a power of two >= 2 (capacity + 1))
};

enum WalkStatus : int32_t {
    kWalkDone = 0,        // the walk ran to its end (queue empty or limit counted)
    kWalkNeed = 1,        // a run reached the end of a stream cut short (open)
    kWalkBroken = 2,      // a score outside [0, max_score] or an entry missing
                          // from the table: inconsistent input, reported
};

struct WalkResult {
    int32_t status;
    int32_t n_hits;
    int32_t counted;
    int64_t pops, runs;
};

TPZ_HD uint32_t walk_table_size(int32_t capacity) {
    uint32_t t = 16;
    while (t < 2u * (uint32_t)(capacity + 1) + 2u) t <<= 1;
    return t;
}

TPZ_HD uint32_t walk_mix(uint32_t k) {
    k ^= k >> 16;
    k *= 0x7feb352dU;
    k ^= k >> 15;
    k *= 0x846ca68bU;
    k ^= k >> 16;
    return k;
}

// The table is at most half full, so every probe sequence meets an empty
// slot; the loops are bounded by the table size all the same.
TPZ_HD int32_t table_find(const WalkMem& m, uint32_t k) {
    uint32_t h = walk_mix(k) & m.table_mask;
    for (uint32_t n = 0; n <= m.table_mask; ++n, h = (h + 1) & m.table_mask) {
        if (m.val[h] < 0) return -1;
        if (m.key[h] == k) return m.val[h];
    }
    return -1;
}

TPZ_HD void table_put(const WalkMem& m, uint32_t k, int32_t e) {
    uint32_t h = walk_mix(k) & m.table_mask;
    for (uint32_t n = 0; n < m.table_mask && m.val[h] >= 0; ++n) h = (h + 1) & m.table_mask;
    m.key[h] = k;
    m.val[h] = e;
}

// Backward-shift deletion keeps every probe sequence unbroken.  False when
// k is not in the table.
TPZ_HD bool table_erase(const WalkMem& m, uint32_t k) {
    uint32_t h = walk_mix(k) & m.table_mask;
    for (uint32_t n = 0;; ++n) {
        if (n > m.table_mask || m.val[h] < 0) return false;
        if (m.key[h] == k) break;
        h = (h + 1) & m.table_mask;
    }
    uint32_t j = h;
    for (uint32_t moved = 0; moved <= m.table_mask; ++moved) {
        m.val[h] = -1;
        for (;;) {
            j = (j + 1) & m.table_mask;
            if (m.val[j] < 0) return true;
            const uint32_t home = walk_mix(m.key[j]) & m.table_mask;
            // the entry at j may stay where its home lies cyclically in (h, j]
            const bool stays = h <= j ? (h < home && home <= j) : (h < home || home <= j);
            if (!stays) break;
        }
        m.key[h] = m.key[j];
        m.val[h] = m.val[j];
        h = j;
    }
    return true;
}

TPZ_HD void link_tail(const WalkMem& m, int32_t e, int32_t& min_score) {
    const int32_t s = m.entry[e].score;
    m.entry[e].next = -1;
    m.entry[e].prev = m.tail[s];
    if (m.tail[s] >= 0) m.entry[m.tail[s]].next = e;
    else m.head[s] = e;
    m.tail[s] = e;
    if (s < min_score) min_score = s;
}

TPZ_HD void unlink(const WalkMem& m, int32_t e) {
    const int32_t s = m.entry[e].score;
    const int32_t p = m.entry[e].prev, n = m.entry[e].next;
    if (p >= 0) m.entry[p].next = n;
    else m.head[s] = n;
    if (n >= 0) m.entry[n].prev = p;
    else m.tail[s] = p;
}

// Element scans of a run: from element i, the first j >= i that ends the
// run (j == len, or j past the run's first element holding kStreamEnd or a
// value below the leader) or that the hit list must see one by one (every
// element while the list is not full, else one above its worst score).  The
// elements before j each count without changing the list.
struct SerialScan {
    TPZ_HD int32_t operator()(const int16_t* v, int32_t i, int32_t len, int32_t start,
                              int32_t leader, bool all_slow, int32_t worst) const {
        for (int32_t j = i;; ++j) {
            if (j >= len) return j;
            const int32_t x = v[j];
            if (j > start && (x == kStreamEnd || x < leader)) return j;
            if (all_slow || x > worst) return j;
        }
    }
};

TPZ_HD bool run_ends_at(const int16_t* v, int32_t j, int32_t len, int32_t start,
                        int32_t leader) {
    return j >= len || (j > start && (v[j] == kStreamEnd || v[j] < leader));
}

// hitlist_add (hit.c:126-193) of an element that is not simply counted:
// the list is not full, or the element beats its worst score.  Returns
// whether the reference returns non-NULL (the walk counts it): a new protein.
TPZ_HD bool hit_add_slow(const WalkMem& m, int32_t capacity, int32_t max_score, int32_t x,
                         uint32_t idx, int32_t& size, int32_t& n_entries, int32_t& n_free,
                         int32_t& min_score, bool& broken) {
    if (x < 0 || x > max_score) {           // outside the score lists
        broken = true;
        return false;
    }
    const int32_t e = table_find(m, idx);
    if (e < 0) {
        const int32_t ne = n_free > 0 ? m.free_stack[--n_free] : n_entries++;
        m.entry[ne].db_index = idx;
        m.entry[ne].score = x;
        link_tail(m, ne, min_score);
        table_put(m, idx, ne);
        if (++size > capacity) {
            // evict the worst: the oldest of the lowest score
            while (min_score <= max_score && m.head[min_score] < 0) ++min_score;
            const int32_t w = m.head[min_score];
            unlink(m, w);
            if (!table_erase(m, m.entry[w].db_index)) broken = true;
            m.free_stack[n_free++] = w;
            --size;
        }
        return true;
    }
    if (m.entry[e].score < x) {
        unlink(m, e);
        m.entry[e].score = x;
        link_tail(m, e, min_score);
    }
    return false;
}

// The walk of one query.  Returns the hits in the priority tree's traversal
// order (score descending, newest first among equals; rbtree.c:646-669)
// through out_index / out_score (room for capacity entries).  Scan finds the
// elements that count without touching the list (SerialScan on the host, a
// wavefront-wide scan on the GPU: both give the same j).
// Mem is WalkMem, or a type derived from it whose `top` lives elsewhere (the
// GPU walk keeps the stack tops in LDS).
template <typename Scan, typename Mem>
TPZ_HD WalkResult walk_query(const WalkStreams& st, const Mem& m, int32_t capacity,
                             int32_t limit, int32_t max_score, const Scan& scan,
                             uint32_t* out_index, int32_t* out_score) {
    WalkResult res{kWalkDone, 0, 0, 0, 0};
    // the streams in creation order, each with its first run (element 0)
    int32_t queued = 0;
    for (int32_t s = 0; s < st.n; ++s) {
        if (st.len[s] <= 0) continue;
        const int32_t x = st.vals[st.base[s]];
        if (x < 0 || x > max_score) {
            res.status = kWalkBroken;
            return res;
        }
        m.node[s].start = 0;
        m.node[s].next = m.top[x];
        m.top[x] = s;
        ++queued;
    }
    int32_t cur = max_score;           // no run waits above it
    int32_t size = 0, n_entries = 0, n_free = 0, min_score = max_score + 1;
    bool stop = false, broken = false;
    while (queued > 0 && !stop) {
        while (m.top[cur] < 0) --cur;
        const int32_t stream = m.top[cur];
        const int32_t leader = cur, start = m.node[stream].start;
        m.top[cur] = m.node[stream].next;
        --queued;
        ++res.runs;
        const int16_t* v = st.vals + st.base[stream];
        const uint32_t* pr = st.prot + st.base[stream];
        const int32_t len = st.len[stream];
        int32_t i = start;
        for (;;) {
            const bool full = size >= capacity;
            if (full)
                while (min_score <= max_score && m.head[min_score] < 0) ++min_score;
            // with the list full, an element counts at once when the worst
            // hit scores >= it (a list of capacity 0 has no worst hit and
            // counts everything: the reference dereferences NULL there)
            const int32_t worst = size == 0 ? 0x7fffffff : min_score;
            const int32_t j = scan(v, i, len, start, leader, !full, worst);
            const int32_t n_fast = j - i;
            if (n_fast > 0) {
                if (res.counted < limit && res.counted + n_fast >= limit) {
                    res.pops += limit - res.counted;
                    res.counted = limit;
                    stop = true;
                    break;
                }
                res.counted += n_fast;
                res.pops += n_fast;
            }
            if (run_ends_at(v, j, len, start, leader)) {
                if (j < len && v[j] != kStreamEnd) {
                    // the stream's next run, on top of its leader's stack
                    // (v[j] < leader)
                    if (v[j] < 0) {
                        res.status = kWalkBroken;
                        stop = true;
                        break;
                    }
                    m.node[stream].start = j;
                    m.node[stream].next = m.top[v[j]];
                    m.top[v[j]] = stream;
                    ++queued;
                } else if (j >= len && st.open && st.open[stream]) {
                    res.status = kWalkNeed;
                    stop = true;
                }
                break;
            }
            ++res.pops;
            const bool counts = hit_add_slow(m, capacity, max_score, v[j], pr[j], size,
                                             n_entries, n_free, min_score, broken);
            if (broken) {
                res.status = kWalkBroken;
                stop = true;
                break;
            }
            if (counts && ++res.counted == limit) {
                stop = true;
                break;
            }
            i = j + 1;
        }
    }
    int32_t n = 0;
    for (int32_t s = max_score; s >= 0 && n < size; --s)
        for (int32_t e = m.tail[s]; e >= 0 && n < size; e = m.entry[e].prev) {
            out_index[n] = m.entry[e].db_index;
            out_score[n] = s;
            ++n;
        }
    res.n_hits = n;
    return res;
}

// Bytes of one query's WalkMem for n_streams streams, and where each array
// starts (256-byte aligned), from `base`.
struct WalkMemLayout {
    size_t top, node, entry, free_stack, head, tail, key, val, total;
    uint32_t table;
};

TPZ_HD size_t walk_align(size_t x) { return (x + 255) & ~(size_t)255; }

TPZ_HD WalkMemLayout walk_mem_layout(int32_t n_streams, int32_t capacity, int32_t max_score) {
    WalkMemLayout L;
    L.table = walk_table_size(capacity);
    size_t o = 0;
    L.top = o; o = walk_align(o + 4 * (size_t)(max_score + 1));
    L.head = o; o = walk_align(o + 4 * (size_t)(max_score + 1));
    L.tail = o; o = walk_align(o + 4 * (size_t)(max_score + 1));
    L.val = o; o = walk_align(o + 4 * (size_t)L.table);
    // the arrays above start as -1; those below need no initial value
    L.key = o; o = walk_align(o + 4 * (size_t)L.table);
    L.node = o; o = walk_align(o + sizeof(WalkNode) * (size_t)(n_streams + 1));
    L.entry = o; o = walk_align(o + sizeof(WalkEntry) * (size_t)(capacity + 1));
    L.free_stack = o; o = walk_align(o + 4 * (size_t)(capacity + 1));
    L.total = o;
    return L;
}

// Bytes at the start of a WalkMem block that must hold -1 (all bits set).
TPZ_HD size_t walk_mem_init_bytes(const WalkMemLayout& L) { return L.key; }

TPZ_HD WalkMem walk_mem_at(char* base, const WalkMemLayout& L) {
    WalkMem m;
    m.top = reinterpret_cast<int32_t*>(base + L.top);
    m.node = reinterpret_cast<WalkNode*>(base + L.node);
    m.entry = reinterpret_cast<WalkEntry*>(base + L.entry);
    m.free_stack = reinterpret_cast<int32_t*>(base + L.free_stack);
    m.head = reinterpret_cast<int32_t*>(base + L.head);
    m.tail = reinterpret_cast<int32_t*>(base + L.tail);
    m.key = reinterpret_cast<uint32_t*>(base + L.key);
    m.val = reinterpret_cast<int32_t*>(base + L.val);
    m.table_mask = L.table - 1;
    return m;
}

}  // namespace tpz

#endif
