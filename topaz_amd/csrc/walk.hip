// The seed walk of one query on the host (tpz_walk_query): every stream
// extended to its end through the caller's callback, then walk_core.h's
// walk, the code the GPU runs per query in the batched search.
#include "walk.h"

#include <string.h>

#include <algorithm>

#include "../../include/topaz_amd.h"
#include "runtime.h"

namespace tpz {

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
        in->max_score > 32767 || in->capacity < 0)
        return fail("tpz_walk_query: bad sizes");
    ProteinIndex proteins;
    if (!proteins.build(in->t_size, in->n_proteins, in->protein_start, in->protein_len))
        return fail("tpz_walk_query: the proteins do not tile the text");
    // every stream in creation order (query.c:104-131: per offset the forward
    // suffix at pos, then the backward one at pos - 1), up to its first
    // non-seed
    std::vector<int64_t> base;
    std::vector<int32_t> len;
    std::vector<int16_t> vals, buf;
    std::vector<uint32_t> prot;
    int64_t refills = 0, elements = 0;
    const int64_t t_size = (int64_t)in->t_size;
    for (int32_t k = 0; k < in->n_offsets; ++k) {
        const int64_t pos = in->positions[k];
        if (pos == -1) continue;
        for (int d = 0; d < 2; ++d) {
            const int32_t dir = d ? -1 : 1;
            const int64_t first = d ? pos - 1 : pos;
            if (d && pos == 0) continue;
            if (first < 0 || first >= t_size) continue;       // no suffix there
            base.push_back((int64_t)vals.size());
            int64_t have = 0;
            int32_t n = in->first_chunk;
            for (bool more = true; more;) {
                buf.assign((size_t)n, kStreamEnd);
                if (extend(ctx, in->offsets[k], first + have * dir, dir, n, buf.data()))
                    return fail("tpz_walk_query: the extension callback failed");
                if (have > 0) ++refills;
                elements += n;
                int32_t take = 0;
                while (take < n && buf[take] != kStreamEnd) ++take;
                for (int32_t i = 0; i < take; ++i) {
                    const int64_t slot = first + (have + i) * dir;
                    vals.push_back(buf[i]);
                    prot.push_back(proteins.of((uint64_t)in->sa[slot]));
                }
                have += take;
                more = take == n;
                n = have == take ? in->refill : (int32_t)std::min<int64_t>(2LL * n, 1 << 20);
            }
            len.push_back((int32_t)have);
        }
    }
    const WalkMemLayout L = walk_mem_layout((int32_t)len.size(), in->capacity, in->max_score);
    std::vector<uint64_t> mem_words(L.total / 8 + 1);
    char* mem = reinterpret_cast<char*>(mem_words.data());
    memset(mem, 0xff, walk_mem_init_bytes(L));
    const WalkMem m = walk_mem_at(mem, L);
    const WalkStreams st{(int32_t)len.size(), base.data(), len.data(), nullptr, vals.data(),
                         prot.data()};
    std::vector<uint32_t> ix((size_t)in->capacity + 1);
    std::vector<int32_t> sc((size_t)in->capacity + 1);
    const WalkResult r = walk_query(st, m, in->capacity, in->number_of_seeds, in->max_score,
                                    SerialScan{}, ix.data(), sc.data());
    if (r.status != kWalkDone) return fail("tpz_walk_query: the hit list lost an entry");
    const int32_t cap = *n_hits;
    *n_hits = r.n_hits;
    for (int32_t i = 0; i < r.n_hits && i < cap; ++i) {
        if (db_index) db_index[i] = ix[i];
        if (score) score[i] = sc[i];
    }
    if (stats) {
        stats->pops = r.pops;
        stats->runs = r.runs;
        stats->refills = refills;
        stats->elements = elements;
        stats->counted = r.counted;
    }
    return 0;
}
