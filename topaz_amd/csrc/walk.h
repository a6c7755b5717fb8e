// Host pieces of the seed walk: the protein of a text position, on the
// host and as the tables the GPU reads.  The walk itself is walk_core.h.
#ifndef TOPAZ_AMD_WALK_H
#define TOPAZ_AMD_WALK_H

#include <stdint.h>

#include <vector>

#include "walk_core.h"

namespace tpz {

// db_index (db.c:176-205) over a text that TOPAZ's index tiles: each
// protein's letters and its '\n', back to back from position 0.  The
// protein holding position p is the one holding p & ~63 or one of the next.
class ProteinIndex {
public:
    // false when the proteins do not tile [0, t_size)
    bool build(uint64_t t_size, uint64_t n_proteins, const uint64_t* start, const int32_t* len);
    uint32_t of(uint64_t pos) const {
        uint64_t k = block_[pos >> 6];
        while (pos > end_[k]) ++k;
        return (uint32_t)k;
    }
    uint64_t size() const { return end_.size(); }
    const std::vector<uint32_t>& blocks() const { return block_; }
    const std::vector<uint64_t>& ends() const { return end_; }

private:
    std::vector<uint32_t> block_;
    std::vector<uint64_t> end_;      // position of each protein's '\n'
};

}  // namespace tpz

#endif
