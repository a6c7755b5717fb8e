// Kernel-side argument block of the batched gapless extension kernel.
#ifndef TOPAZ_AMD_GAPLESS_KERNELS_H
#define TOPAZ_AMD_GAPLESS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// One seed: query number, offset in the query, and a text position (or a
// suffix-array slot when GaplessArgs::sa is set).  16 bytes.
struct SeedRec {
    int32_t query;
    int32_t offset;
    int64_t position;
};

// Letter codes.  Every byte value present in the text (other than '\n') has
// a text code 0..30 and '\n' has kTextTerm.  A query byte that occurs in the
// text has the same code, any other query byte a code 32 + its matrix row,
// so two codes are equal exactly when the two bytes are.  kQueryTerm ends a
// query.  Query codes are stored doubled: the 16-bit pair (text code << 8 |
// 2 * query code) is then the byte address of W(text code, query code) in a
// 32 x 128 table of int16.
constexpr int kTextTerm = 31;
constexpr int kQueryTerm = 63;
constexpr int kForeignCode = 32;
constexpr int kTableRows = 32;
constexpr int kTableCols = 128;
constexpr int kTermScore = -30000;   // past a terminator no prefix sum is >= 0 again
constexpr int kWindow = 128;         // sa_hsp's window on each side (search.c:99-190)
constexpr int kTextPad = 256;        // terminator bytes after each code array
constexpr int kHeadBytes = 2 * kWindow + kTextPad;

struct GaplessArgs {
    const uint8_t* t_fwd;       // text codes, t_size bytes then kTextPad x kTextTerm
    const uint8_t* t_head;      // kWindow x kTextTerm, then the codes of T[0 .. kWindow) with
                                // T[0] as kTextTerm, then terminators: the left window of
                                // idx < kWindow
    int64_t t_size;
    const int64_t* sa;          // may be null: positions are text positions
    const uint8_t* q_fwd;       // per query: doubled codes of q[0..m-1], then 2 * kQueryTerm
    const int64_t* qf_off;
    const uint8_t* q_rev;       // per query: doubled codes of q[m-1..1], then two terminators
    const int64_t* qr_off;
    const int32_t* q_len;
    int32_t n_queries;          // seeds naming a query outside [0, n_queries) get length -1
    const SeedRec* seeds;
    int64_t n_seeds;
    const int16_t* table;       // kTableRows x kTableCols, W(text code, query code)
    int32_t* out_score;         // 0 where the length is -1
    int32_t* out_length;
};

// sa_search (search.c:237-361) for offsets of one query: the suffix-array
// interval of the query suffix at each offset.
struct SaSearchArgs {
    const char* text;           // db->T as bytes
    int64_t t_size;
    const int64_t* sa;          // db->SA
    const char* query;          // the query's bytes
    int64_t q_len;
    const int32_t* offsets;
    int32_t n;
    int64_t* out;               // per offset: idx, num_matches, last_position
    int32_t* status;            // per offset: sa_search's return value
};
constexpr int kSaSearchQueryLds = 32768;   // a longer query is read from memory

// Where query_find_suffixes starts its suffixes (query.c:104-131): for each
// (query, offset) record, the number of suffixes of the text that sort
// below the query's suffix at that offset -- the index sa_search reports,
// whether or not the suffix occurs.
struct LowerBoundArgs {
    const char* text;           // db->T
    int64_t t_size;
    const int64_t* sa;          // db->SA
    const uint32_t* sa32;       // when set, db->SA as uint32, read instead of sa
    const char* q_chars;        // the queries' bytes, back to back
    const int64_t* q_off;
    const int32_t* q_len;
    const int32_t* rec_query;   // per record: query number and offset
    const int32_t* rec_offset;
    int64_t n;
    int64_t* out;               // per record: the index
};

// A walk stream (query.c:112-128): the suffix-array slots first,
// first + dir, ... of one (query, offset).  24 bytes.
struct StreamDef {
    int32_t query;
    int32_t offset;
    int64_t first;
    int32_t dir;                // +1 or -1
    int32_t pad;
};

// One round of stream extension: request r takes the `chunk` slots of
// stream req_stream[r] from its element `have` on (every listed stream
// holds `have` elements), extended with sa_hsp.  Outputs per request, at
// r * chunk: the score where the extension is a seed (length >= min_len),
// -1 otherwise (also for a slot outside the suffix array), and the protein
// (db_index) of the slot's text position from a table of every 64th
// position's protein and each protein's last position (its newline;
// 0xffffffff outside the suffix array); and req_end[r], lowered to the
// first element holding -1 (preset to chunk by the caller).
struct StreamRoundArgs {
    GaplessArgs g;              // text, suffix array (required), queries, table
    const StreamDef* defs;
    const int32_t* req_stream;
    int64_t n_req;
    int64_t have;
    int32_t chunk;
    int32_t min_len;
    int16_t* out;
    uint32_t* out_protein;
    int32_t* req_end;
    const uint32_t* prot_block;
    const uint64_t* prot_end;
    const uint32_t* sa32;       // when set, the suffix array as uint32 (a text under 4 GiB),
                                // read instead of g.sa
};

}  // namespace tpz

extern "C" hipError_t tpz_launch_gapless(const tpz::GaplessArgs* args, int grid,
                                         hipStream_t stream);
extern "C" hipError_t tpz_launch_sa_search(const tpz::SaSearchArgs* args, hipStream_t stream);
extern "C" hipError_t tpz_launch_lower_bound(const tpz::LowerBoundArgs* args, hipStream_t stream);
extern "C" hipError_t tpz_launch_stream_round(const tpz::StreamRoundArgs* args, int grid,
                                              hipStream_t stream);

#endif
