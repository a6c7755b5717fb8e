// Internal host-side helpers shared by the library's translation units.
#ifndef TOPAZ_AMD_RUNTIME_H
#define TOPAZ_AMD_RUNTIME_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <utility>
#include <functional>
#include <vector>

struct tpz_server;

namespace tpz {

int fail(const char* fmt, ...);          // set the error string, return -1
int check(hipError_t e, const char* what);
int ensure_init();                       // tpz_init(0) on first use
hipStream_t stream();
int num_cu();
int device_id();                         // the device tpz_init selected

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    void reset();
    int alloc(size_t n);
    int upload(const void* src, size_t n);
    // alloc(n), then the bytes through pinned staging (stage_to_device)
    int upload_large(const void* src, size_t n);
};

// Bytes [0, n) of device memory dst, produced on the host by fill(out,
// begin, end) (which writes source bytes [begin, end) to out), through pinned
// staging buffers: host threads fill one buffer while the copy engine drains
// the previous ones.  For the multi-gigabyte database arrays, whose pageable
// copies and single-threaded conversions would otherwise dominate the
// upload.  Returns when every byte is on the device.
int stage_to_device(void* dst, size_t n,
                    const std::function<void(uint8_t* out, size_t begin, size_t end)>& fill);

// ---- batched Smith-Waterman planning (shared by batches and the server) ----
constexpr int kGroupsPerCu = 2;     // default resident sw_forward workgroups per CU
constexpr int kGroupWaves = 4;      // default wavefronts per sw_forward workgroup (4 or 8)
constexpr int kTaskPairs = 128;     // two pairs per lane of a wavefront
constexpr int kTaskQueries = 4;     // queries whose profiles a wavefront can hold

// Longest query or subject of the forward kernel: its keys hold columns and
// rows below 2^24 (with room for a run's padding columns).
constexpr int32_t kMaxSwLength = (1 << 24) - 64;

// Columns a pair with an n-residue subject takes in a lane half's run (the
// forward kernel's item_cols): n rounded up to 4, at least 8.
inline int32_t item_columns(int32_t n) { return std::max<int32_t>(8, (n + 3) & ~3); }

// Tasks of the kernel: pairs of up to 4 queries.  A task's first (up to 128)
// entries start the runs of its wavefront's lane halves; next[] chains the
// further entries of a run (-1 ends it).
struct Tasks {
    std::vector<int32_t> index;   // pair numbers, grouped per task (one entry each)
    std::vector<int32_t> next;    // per entry: the next entry of its run, -1 = none
    std::vector<int32_t> begin;   // task boundaries into index
    std::vector<int32_t> query;   // kTaskQueries query numbers per task, -1 = none
    std::vector<int32_t> cols;    // per task: columns of its longest run
    std::vector<int64_t> cost;    // rows (whole chunks) x cols
    void clear() { index.clear(); next.clear(); begin.clear(); query.clear(); cols.clear(); cost.clear(); }
    void open(const int32_t* qs, int nq) {
        begin.push_back((int32_t)index.size());
        for (int k = 0; k < kTaskQueries; ++k) query.push_back(k < nq ? qs[k] : -1);
        cols.push_back(0);
        cost.push_back(0);
    }
    int32_t count() const { return (int32_t)cost.size(); }
};

// Per pair: its subject's offset / 16 (s_off 16-byte aligned) and length,
// gathered once so that planning reads the subject tables at random once per
// pair.
std::vector<int2> pair_subjects(const int32_t* pair_subject, int64_t n_pairs, const int64_t* s_off,
                                const int32_t* s_len);
// Per entry of t: the forward kernel's record (SwKernelArgs::entry_rec):
// subject offset / 16, subject length (both from pair_sub), next entry,
// query slot within the entry's task.
void build_entry_records(const Tasks& t, const int32_t* pair_query, const int2* pair_sub,
                         std::vector<int4>& rec);

// A task set on the device.
struct DeviceTasks {
    Tasks host;
    DeviceBuffer d_index, d_begin, d_query, d_cols, d_rec;
    int upload(const std::vector<int4>& rec) {
        return d_index.upload(host.index.data(), host.index.size() * 4) |
               d_begin.upload(host.begin.data(), host.begin.size() * 4) |
               d_query.upload(host.query.data(), host.query.size() * 4) |
               d_cols.upload(host.cols.data(), host.cols.size() * 4) |
               d_rec.upload(rec.data(), rec.size() * sizeof(int4));
    }
};

// Gap penalties the reference's ssw takes (uint8_t, ssw.h:107-118).
inline bool gaps_valid(int32_t go, int32_t ge) { return go >= 0 && go <= 255 && ge >= 0 && ge <= 255; }
// Whether the reference's lane recurrence is textbook affine-gap local
// alignment (gap_open > gap_extend), which the forward, pair and reverse
// kernels compute; otherwise the lane kernel emulates the lanes.
inline bool textbook_gaps(int32_t go, int32_t ge) { return go > ge; }

// Pairs of each query, longest subject first (ties: lower pair number):
// query q's pairs are pair[begin[q] .. begin[q + 1]), and len[] holds the
// length of each one's subject.
struct QueryPairs {
    std::vector<int64_t> begin;
    std::vector<int32_t> pair, len;
    int64_t count(int32_t q) const { return begin[q + 1] - begin[q]; }
};
QueryPairs pairs_by_query(const std::vector<int32_t>& pair_query, const int2* pair_sub,
                          int32_t n_queries);
// Up to nq_max queries of equal chunk count per task (the persistent
// schedule), the runs of the lane halves filled as described at the
// definition.
void build_packed_tasks(const QueryPairs& per, const std::vector<int32_t>& q_len, int rows,
                        int nq_max, Tasks& t, bool fill = true);
// fn(i) for i in [0, n) on up to `threads` host threads (the caller's
// included); contiguous index ranges per thread.
void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& range_fn);
// fn(i) for every i in [0, n) on up to `threads` host threads (the caller's
// included), each taking the next index from a shared counter.
void parallel_for_each(int64_t n, int threads, const std::function<void(int64_t)>& fn);
// Launches of the server's forward kernel with fewer than `pairs` pairs run
// 8 wavefronts per workgroup (0: every launch runs the configured width).
void server_small_launches(::tpz_server* s, int64_t pairs);
// Host threads the library uses for its parallel loops (TOPAZ_AMD_HOST_THREADS,
// else the machine's, at most 64).
int host_threads();
// Work items of the persistent schedule (`waves` entries each), most
// expensive first.
void build_items(const Tasks& t, int n_slots, double big_fraction, int waves,
                 std::vector<int32_t>& items, int32_t& n_big);

// Letter codes of a text for the gapless kernel (see gapless_kernels.h).
struct TextCodes {
    uint8_t code[256];        // text byte -> code; 0xff: the byte is not in the text
    uint8_t byte_of[32];      // code -> byte
    bool used[32];            // codes of letters of the text (kTextTerm, '\n', excluded)
};

// Query copies of a gapless batch in the kernel's layout: forward and
// reversed doubled codes with terminators, and their offsets.
struct QueryCodes {
    std::vector<uint8_t> fwd, rev;
    std::vector<int64_t> fwd_off, rev_off;
};
int gapless_query_codes(const TextCodes& tc, int32_t n_queries, const char* q_chars,
                        const int64_t* q_off, const int32_t* q_len, QueryCodes& out);
// W(text code, query code) for the kernel, kTableRows x kTableCols int16.
void gapless_table(const TextCodes& tc, const int8_t* matrix, int16_t* table);
// Matrix rows of text bytes and of query bytes (lexicographical.h:19-74),
// 128 entries each; bytes >= 128 use their low seven bits.
void alphabet_rows(int8_t* text_rows, int8_t* query_rows);

// ---- traceback statistics (shared by tpz_stats_batch and the drop-in ABI) ----
// Blocks handed out in order and reused after reset(); never freed here.
struct BlockList {
    std::vector<std::pair<char*, size_t>> blocks;
    size_t cur = 0, off = 0;
    void reset() { cur = 0; off = 0; }
    // bytes (rounded up to 256) from the current block, or from a new one of
    // at least twice the last block's size got from grow(size, ctx)
    char* take(size_t bytes, void* (*grow)(size_t, void*), void* ctx);
};

// Memory and stream of one statistics run: device and pinned host memory
// valid until the next reset(), the stream the work goes to, and a wait for
// the work queued on it so far.  Failures return nullptr / -1 with the
// error string set.
struct RunMem {
    virtual ~RunMem() = default;
    virtual void reset() = 0;
    virtual void* dev(size_t bytes) = 0;
    virtual char* host(size_t bytes) = 0;
    virtual hipStream_t stream() = 0;
    virtual int wait() = 0;
};

struct StatsInput {
    int32_t n_queries;
    const int8_t* q_res;
    const int64_t* q_off;
    const int32_t* q_len;
    int32_t n_subjects;
    const int8_t* s_res;
    const int64_t* s_off;
    const int32_t* s_len;
    int32_t n_hits;
    const int32_t *hit_query, *hit_subject, *read_end1, *ref_end1, *score1;
    const int8_t* matrix;
    int32_t gap_open, gap_extend;
};
struct StatsOutput {
    // per hit: status, begins (0-based), and the counts of the traceback;
    // ident is the identical count, or -1 / -2 for the size guard / a
    // failed traceback
    std::vector<int32_t> rc, read_begin, ref_begin, len, gapo, mis, ident;
};
int stats_check(const StatsInput& in);
int stats_compute(const StatsInput& in, RunMem& m, StatsOutput& out);
double stats_identity(int32_t ident, int32_t length);

}  // namespace tpz

// Device-resident text of the gapless kernel.
struct tpz_text {
    uint64_t t_size = 0;
    bool has_sa = false;
    tpz::TextCodes codes;
    tpz::DeviceBuffer d_fwd, d_head, d_sa;
};

#endif
