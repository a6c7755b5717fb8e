// The seed walk of a query and its hit list, on the host: TOPAZ's priority
// queue over suffix-array streams (priority_queue.c:36-56, suffix.c:40-65)
// and hitlist_add (hit.c:126-193), reproduced exactly, over stream elements
// computed in bulk on the GPU.
//
// A stream is one suffix of the reference's walk: query offset o, a first
// suffix-array slot and a direction; element k is the gapless extension at
// slot first + k * direction.  The walk pops the stream with the largest
// current score (ties: the one inserted last, rbtree.c:72-91, 410-425),
// hands the element to the hit list, and puts the stream back with its next
// element while that one is a seed.
//
// Blocks.  After a stream's element x is popped (x is then the largest score
// in the queue), its next element y >= x is at once the largest again, and
// the newest among equals, so it is popped next; the same holds for every
// following element >= x.  The reference therefore pops a stream in runs:
// a run starts at an element smaller than every earlier element of its
// stream (a strict prefix minimum, the run's leader) and takes every
// following element >= the leader.  Runs of one stream have strictly
// decreasing leaders.  Popping whole runs from a queue keyed by (leader,
// insertion number), a run being inserted when the run before it in its
// stream ends, gives the reference's pop sequence element for element, with
// one queue operation per run instead of per element.
//
// The walk is resumable: when a run reaches the end of its stream's computed
// elements, run() returns kNeed and the caller supplies more (the batched
// search gathers the requests of many queries into one GPU launch).
#ifndef TOPAZ_AMD_WALK_H
#define TOPAZ_AMD_WALK_H

#include <stdint.h>

#include <algorithm>
#include <vector>

namespace tpz {

// Element value of a stream slot: the gapless score when its extension is a
// seed (length >= min_hit_length), else kStreamEnd (also for a slot outside
// the suffix array); a walk never continues past kStreamEnd.
constexpr int16_t kStreamEnd = -1;

// The hit list of one query (hit.c:99-193): at most `capacity` proteins, each
// with its best gapless score.  Order is (score, insertion number): the
// priority tree's in-order with equal keys in insertion order, a re-put
// counting as a new insertion (rbtree.c:496-505).  Proteins are found by
// db_index of the element's text position, given by the caller.
class HitList {
public:
    void reset(int32_t capacity, int32_t max_score);
    // hitlist_add: true where the reference returns non-NULL (the walk counts
    // it), i.e. a new protein, or the list full and its worst score >= score.
    // protein(): the db_index of the element (asked only when needed).
    template <typename ProteinOf>
    bool add(int32_t score, ProteinOf&& protein);
    int32_t size() const { return size_; }
    // The hits in the reference's traversal order (score descending, newest
    // first among equal scores, rbtree.c:646-669).
    void traversal(std::vector<uint64_t>& db_index, std::vector<int32_t>& score) const;

private:
    struct Entry {
        uint64_t db_index;
        int32_t score;
        int32_t prev, next;      // bucket list of its score, oldest first
    };
    int32_t capacity_ = 0, size_ = 0, min_score_ = 0, max_score_ = 0;
    std::vector<Entry> entry_;
    std::vector<int32_t> free_;
    std::vector<int32_t> head_, tail_;        // per score: oldest / newest entry
    // open-addressing map db_index -> entry (linear probing, backward-shift delete)
    std::vector<uint64_t> key_;
    std::vector<int32_t> val_;
    uint64_t mask_ = 0;

    int32_t find(uint64_t k) const;
    void map_put(uint64_t k, int32_t e);
    void map_erase(uint64_t k);
    void link_tail(int32_t e);
    void unlink(int32_t e);
    int32_t min_entry();
};

// db_index (db.c:176-205) over a text that TOPAZ's index tiles: each
// protein's letters and its '\n', back to back from position 0.  The
// protein holding position p is the one holding p & ~63 or one of the next.
class ProteinIndex {
public:
    // false when the proteins do not tile [0, t_size)
    bool build(uint64_t t_size, uint64_t n_proteins, const uint64_t* start, const int32_t* len);
    uint64_t of(uint64_t pos) const {
        uint64_t k = block_[pos >> 6];
        while (pos > end_[k]) ++k;
        return k;
    }
    uint64_t size() const { return end_.size(); }
    const std::vector<uint32_t>& blocks() const { return block_; }
    const std::vector<uint64_t>& ends() const { return end_; }

private:
    std::vector<uint32_t> block_;
    std::vector<uint64_t> end_;      // position of each protein's '\n'
};

struct WalkStats {
    int64_t pops = 0;         // elements popped
    int64_t runs = 0;         // runs popped
    int64_t refills = 0;      // times the walk waited for elements
    int64_t elements = 0;     // elements supplied (computed)
};

class Walk {
public:
    enum Status { kDone = 0, kNeed = 1 };

    // number_of_seeds: the walk stops once the hit list has counted this many
    // (query.c:144-150).
    void reset(int32_t number_of_seeds, int32_t capacity, int32_t max_score);
    // A stream in creation order (query.c:104-131: per offset the forward
    // suffix at pos, then the backward one at pos - 1), with its computed
    // elements 0 .. n-1; element 0 must be a seed.  `complete`: no element
    // after these exists (the suffix array ends).
    void add_stream(int32_t offset, int64_t first_slot, int32_t dir, const int16_t* vals,
                    int32_t n, bool complete);
    // The same, taking over computed elements that end at their first
    // kStreamEnd, if any, and each element's protein (db_index) alongside.
    void add_stream(int32_t offset, int64_t first_slot, int32_t dir,
                    std::vector<int16_t>&& vals, std::vector<uint32_t>&& proteins,
                    bool complete);
    // Run until done or until a stream needs more elements.
    // protein_of(stream, element, slot): the db_index of that element's text
    // position, asked only when the hit list needs it.
    template <typename ProteinOf>
    Status run(ProteinOf&& protein_of);
    // The protein of element i of stream s when the stream was given them.
    uint32_t protein(int32_t s, int32_t i) const { return streams_[s].proteins[i]; }
    // The stream run() waits for, and the slot of its first missing element.
    int32_t waiting() const { return cur_stream_; }
    int64_t next_slot(int32_t s) const;
    int32_t offset_of(int32_t s) const { return streams_[s].offset; }
    int32_t dir_of(int32_t s) const { return streams_[s].dir; }
    int32_t computed(int32_t s) const { return (int32_t)streams_[s].vals.size(); }
    // Append elements to stream s (those after its last computed one), with
    // their proteins when the stream holds them (else proteins is null).
    void supply(int32_t s, const int16_t* vals, const uint32_t* proteins, int32_t n,
                bool complete);
    // Streams whose next unpopped run starts within `margin` elements of the
    // end of their computed elements, best first, at most `limit` of them
    // (the waiting stream excluded): prefetch candidates.
    void hungry(int32_t limit, int32_t margin, std::vector<int32_t>& out) const;

    const HitList& hits() const { return hits_; }
    int32_t counted() const { return counter_; }
    const WalkStats& stats() const { return stats_; }

private:
    struct Stream {
        int32_t offset;
        int32_t dir;
        int64_t first;
        bool complete;
        std::vector<int16_t> vals;
        std::vector<uint32_t> proteins;     // empty: asked of protein_of
    };
    struct Run {
        int16_t leader;
        uint64_t seq;
        int32_t stream;
        int32_t start;
    };
    static bool below(const Run& a, const Run& b) {
        return a.leader != b.leader ? a.leader < b.leader : a.seq < b.seq;
    }
    std::vector<Stream> streams_;
    std::vector<Run> heap_;
    uint64_t seq_ = 0;
    int32_t limit_ = 0, counter_ = 0;
    bool done_ = false;
    // the run being popped
    int32_t cur_stream_ = -1, cur_index_ = 0;     // next element of the run to look at
    int16_t cur_leader_ = 0;
    bool cur_first_ = false;                      // cur_index_ is the run's leader
    HitList hits_;
    WalkStats stats_;

    void push(int16_t leader, int32_t stream, int32_t start);
};

// ---- template bodies --------------------------------------------------------

template <typename ProteinOf>
bool HitList::add(int32_t score, ProteinOf&& protein) {
    if (size_ >= capacity_) {
        // (a list of capacity 0 has no worst hit; the reference dereferences
        // NULL there, hit.c:141-142)
        const int32_t m = min_entry();
        if (m < 0 || entry_[m].score >= score) return true;
    }
    const uint64_t idx = protein();
    const int32_t e = find(idx);
    if (e < 0) {
        int32_t n;
        if (!free_.empty()) {
            n = free_.back();
            free_.pop_back();
        } else {
            n = (int32_t)entry_.size();
            entry_.push_back(Entry{});
        }
        entry_[n].db_index = idx;
        entry_[n].score = score;
        link_tail(n);
        map_put(idx, n);
        ++size_;
        if (size_ > capacity_) {
            const int32_t m = min_entry();
            unlink(m);
            map_erase(entry_[m].db_index);
            free_.push_back(m);
            --size_;
        }
        return true;
    }
    if (entry_[e].score < score) {
        unlink(e);
        entry_[e].score = score;
        link_tail(e);
    }
    return false;
}

template <typename ProteinOf>
Walk::Status Walk::run(ProteinOf&& protein_of) {
    if (done_) return kDone;
    for (;;) {
        if (cur_stream_ < 0) {
            if (heap_.empty()) {
                done_ = true;
                return kDone;
            }
            std::pop_heap(heap_.begin(), heap_.end(), below);
            const Run r = heap_.back();
            heap_.pop_back();
            cur_stream_ = r.stream;
            cur_index_ = r.start;
            cur_leader_ = r.leader;
            cur_first_ = true;
            ++stats_.runs;
        }
        Stream& s = streams_[cur_stream_];
        const int32_t n = (int32_t)s.vals.size();
        const int16_t* v = s.vals.data();
        for (;;) {
            const int32_t i = cur_index_;
            if (i >= n) {
                if (s.complete) {
                    cur_stream_ = -1;
                    break;
                }
                ++stats_.refills;
                return kNeed;
            }
            if (!cur_first_) {
                // the run goes on while elements are >= its leader
                if (v[i] == kStreamEnd) {
                    cur_stream_ = -1;
                    break;
                }
                if (v[i] < cur_leader_) {
                    push(v[i], cur_stream_, i);
                    cur_stream_ = -1;
                    break;
                }
            }
            cur_first_ = false;
            ++stats_.pops;
            cur_index_ = i + 1;
            const int64_t slot = s.first + (int64_t)i * s.dir;
            const int32_t st = cur_stream_;
            if (hits_.add(v[i], [&] { return protein_of(st, i, slot); })) {
                if (++counter_ == limit_) {
                    done_ = true;
                    return kDone;
                }
            }
        }
    }
}

}  // namespace tpz

#endif
