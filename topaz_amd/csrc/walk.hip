// Host side of the seed walk: the hit list and the run queue (walk.h).
#include "walk.h"

#include <string.h>

#include "../../include/topaz_amd.h"
#include "runtime.h"

namespace tpz {

// ---------------------------------------------------------------------------
// hit list
// ---------------------------------------------------------------------------
void HitList::reset(int32_t capacity, int32_t max_score) {
    capacity_ = capacity;
    size_ = 0;
    max_score_ = max_score;
    min_score_ = max_score + 1;
    entry_.clear();
    free_.clear();
    head_.assign((size_t)max_score + 1, -1);
    tail_.assign((size_t)max_score + 1, -1);
    uint64_t cap = 16;
    while (cap < 2 * (uint64_t)(capacity > 0 ? capacity : 1) + 2) cap <<= 1;
    mask_ = cap - 1;
    key_.assign(cap, 0);
    val_.assign(cap, -1);
}

static inline uint64_t mix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
}

int32_t HitList::find(uint64_t k) const {
    for (uint64_t h = mix(k) & mask_;; h = (h + 1) & mask_) {
        if (val_[h] < 0) return -1;
        if (key_[h] == k) return val_[h];
    }
}

void HitList::map_put(uint64_t k, int32_t e) {
    uint64_t h = mix(k) & mask_;
    while (val_[h] >= 0) h = (h + 1) & mask_;
    key_[h] = k;
    val_[h] = e;
}

void HitList::map_erase(uint64_t k) {
    uint64_t h = mix(k) & mask_;
    while (key_[h] != k || val_[h] < 0) h = (h + 1) & mask_;
    // backward-shift deletion keeps every probe sequence unbroken
    uint64_t j = h;
    for (;;) {
        val_[h] = -1;
        for (;;) {
            j = (j + 1) & mask_;
            if (val_[j] < 0) return;
            const uint64_t home = mix(key_[j]) & mask_;
            // move j to h unless its home lies cyclically in (h, j]
            const bool stays = h <= j ? (h < home && home <= j) : (h < home || home <= j);
            if (!stays) break;
        }
        key_[h] = key_[j];
        val_[h] = val_[j];
        h = j;
    }
}

void HitList::link_tail(int32_t e) {
    const int32_t s = entry_[e].score;
    entry_[e].next = -1;
    entry_[e].prev = tail_[s];
    if (tail_[s] >= 0) entry_[tail_[s]].next = e;
    else head_[s] = e;
    tail_[s] = e;
    if (s < min_score_) min_score_ = s;
}

void HitList::unlink(int32_t e) {
    const int32_t s = entry_[e].score;
    const int32_t p = entry_[e].prev, n = entry_[e].next;
    if (p >= 0) entry_[p].next = n;
    else head_[s] = n;
    if (n >= 0) entry_[n].prev = p;
    else tail_[s] = p;
}

// The oldest entry of the lowest score (rb_getmin / rb_popmin: leftmost).
int32_t HitList::min_entry() {
    while (min_score_ <= max_score_ && head_[min_score_] < 0) ++min_score_;
    return min_score_ <= max_score_ ? head_[min_score_] : -1;
}

void HitList::traversal(std::vector<uint64_t>& db_index, std::vector<int32_t>& score) const {
    db_index.clear();
    score.clear();
    for (int32_t s = max_score_; s >= 0; --s)
        for (int32_t e = tail_[s]; e >= 0; e = entry_[e].prev) {
            db_index.push_back(entry_[e].db_index);
            score.push_back(s);
        }
}

// ---------------------------------------------------------------------------
// proteins
// ---------------------------------------------------------------------------
bool ProteinIndex::build(uint64_t t_size, uint64_t n_proteins, const uint64_t* start,
                         const int32_t* len) {
    block_.clear();
    end_.clear();
    if (n_proteins == 0 || n_proteins >= UINT32_MAX) return false;
    uint64_t next = 0;
    end_.resize(n_proteins);
    for (uint64_t k = 0; k < n_proteins; ++k) {
        if (start[k] != next || len[k] < 0) return false;
        end_[k] = start[k] + (uint64_t)len[k];
        next = end_[k] + 1;
    }
    if (next != t_size) return false;
    block_.resize((t_size + 63) / 64);
    uint64_t k = 0;
    for (uint64_t b = 0; b < block_.size(); ++b) {
        while (b * 64 > end_[k]) ++k;
        block_[b] = (uint32_t)k;
    }
    return true;
}

// ---------------------------------------------------------------------------
// walk
// ---------------------------------------------------------------------------
void Walk::reset(int32_t number_of_seeds, int32_t capacity, int32_t max_score) {
    streams_.clear();
    heap_.clear();
    seq_ = 0;
    limit_ = number_of_seeds;
    counter_ = 0;
    done_ = false;
    cur_stream_ = -1;
    cur_index_ = 0;
    cur_first_ = false;
    hits_.reset(capacity, max_score);
    stats_ = WalkStats{};
}

void Walk::push(int16_t leader, int32_t stream, int32_t start) {
    heap_.push_back(Run{leader, seq_++, stream, start});
    std::push_heap(heap_.begin(), heap_.end(), below);
}

// Elements up to the first kStreamEnd (kept: the walk stops there); true
// when one was found.
static bool append_until_end(std::vector<int16_t>& dst, const int16_t* vals, int32_t n) {
    for (int32_t i = 0; i < n; ++i) {
        dst.push_back(vals[i]);
        if (vals[i] == kStreamEnd) return true;
    }
    return false;
}

void Walk::add_stream(int32_t offset, int64_t first_slot, int32_t dir, const int16_t* vals,
                      int32_t n, bool complete) {
    Stream s;
    s.offset = offset;
    s.dir = dir;
    s.first = first_slot;
    s.complete = append_until_end(s.vals, vals, n) || complete;
    streams_.push_back(std::move(s));
    stats_.elements += n;
    push(vals[0], (int32_t)streams_.size() - 1, 0);
}

void Walk::add_stream(int32_t offset, int64_t first_slot, int32_t dir,
                      std::vector<int16_t>&& vals, std::vector<uint32_t>&& proteins,
                      bool complete) {
    Stream s;
    s.offset = offset;
    s.dir = dir;
    s.first = first_slot;
    s.complete = complete || (!vals.empty() && vals.back() == kStreamEnd);
    stats_.elements += (int64_t)vals.size();
    s.vals = std::move(vals);
    s.proteins = std::move(proteins);
    streams_.push_back(std::move(s));
    push(streams_.back().vals[0], (int32_t)streams_.size() - 1, 0);
}

int64_t Walk::next_slot(int32_t s) const {
    const Stream& st = streams_[s];
    return st.first + (int64_t)st.vals.size() * st.dir;
}

void Walk::supply(int32_t s, const int16_t* vals, const uint32_t* proteins, int32_t n,
                  bool complete) {
    Stream& st = streams_[s];
    if (!st.complete) {
        const size_t before = st.vals.size();
        st.complete = append_until_end(st.vals, vals, n) || complete;
        if (proteins) st.proteins.insert(st.proteins.end(), proteins,
                                          proteins + (st.vals.size() - before));
    }
    stats_.elements += n;
}

void Walk::hungry(int32_t limit, int32_t margin, std::vector<int32_t>& out) const {
    out.clear();
    if (limit <= 0) return;
    // the queue's best runs first: a partial selection of the heap
    std::vector<Run> top(heap_);
    const size_t k = std::min<size_t>(top.size(), (size_t)limit * 4);
    std::partial_sort(top.begin(), top.begin() + k, top.end(),
                      [](const Run& a, const Run& b) { return below(b, a); });
    for (size_t i = 0; i < k && (int32_t)out.size() < limit; ++i) {
        const Stream& st = streams_[top[i].stream];
        if (top[i].stream == cur_stream_ || st.complete) continue;
        if ((int32_t)st.vals.size() - top[i].start < margin) out.push_back(top[i].stream);
    }
}

}  // namespace tpz

// ---------------------------------------------------------------------------
// C ABI: one query's walk over extensions computed by the caller
// ---------------------------------------------------------------------------
using namespace tpz;

extern "C" int tpz_walk_query(const tpz_walk_input* in, tpz_extend_fn extend, void* ctx,
                              uint64_t* db_index, int32_t* score, int32_t* n_hits,
                              tpz_walk_stats* stats) {
    if (!in || !extend || !n_hits) return fail("tpz_walk_query: null argument");
    if (in->n_offsets < 0 || in->first_chunk < 1 || in->refill < 1 || in->max_score < 0 ||
        in->max_score > 32767)
        return fail("tpz_walk_query: bad sizes");
    ProteinIndex proteins;
    if (!proteins.build(in->t_size, in->n_proteins, in->protein_start, in->protein_len))
        return fail("tpz_walk_query: the proteins do not tile the text");
    Walk w;
    w.reset(in->number_of_seeds, in->capacity, in->max_score);
    std::vector<int16_t> buf;
    std::vector<int32_t> want;
    const int64_t t_size = (int64_t)in->t_size;
    auto compute = [&](int32_t offset, int64_t first, int32_t dir, int32_t count) -> int {
        buf.assign((size_t)count, kStreamEnd);
        return extend(ctx, offset, first, dir, count, buf.data());
    };
    for (int32_t k = 0; k < in->n_offsets; ++k) {
        const int64_t pos = in->positions[k];
        if (pos == -1) continue;
        for (int d = 0; d < 2; ++d) {
            const int32_t dir = d ? -1 : 1;
            const int64_t first = d ? pos - 1 : pos;
            if (d && pos == 0) continue;
            if (first < 0 || first >= t_size) continue;       // no suffix there
            if (compute(in->offsets[k], first, dir, in->first_chunk))
                return fail("tpz_walk_query: the extension callback failed");
            if (buf[0] == kStreamEnd) continue;
            w.add_stream(in->offsets[k], first, dir, buf.data(), in->first_chunk, false);
            want.push_back(in->refill);
        }
    }
    const int64_t* sa = in->sa;
    auto protein_of = [&](int32_t, int32_t, int64_t slot) {
        return proteins.of((uint64_t)sa[slot]);
    };
    while (w.run(protein_of) == Walk::kNeed) {
        const int32_t s = w.waiting();
        const int32_t n = want[s];
        want[s] = std::min(2 * n, 1 << 16);
        if (compute(w.offset_of(s), w.next_slot(s), w.dir_of(s), n))
            return fail("tpz_walk_query: the extension callback failed");
        w.supply(s, buf.data(), nullptr, n, false);
    }
    std::vector<uint64_t> ix;
    std::vector<int32_t> sc;
    w.hits().traversal(ix, sc);
    const int32_t cap = *n_hits;
    *n_hits = (int32_t)ix.size();
    for (int32_t i = 0; i < (int32_t)ix.size() && i < cap; ++i) {
        if (db_index) db_index[i] = ix[i];
        if (score) score[i] = sc[i];
    }
    if (stats) {
        stats->pops = w.stats().pops;
        stats->runs = w.stats().runs;
        stats->refills = w.stats().refills;
        stats->elements = w.stats().elements;
        stats->counted = w.counted();
    }
    return 0;
}
