// TOPAZ's search of a batch of queries against a resident database, on one
// MI355X: the path of engine.c:193-248 (_full_search) per query --
// query_find_suffixes, query_fetch_alignment_batch, query_perform_alignments,
// query_perform_statistics, output_top_hits -- with every query of the batch
// going through each stage together.
//
//   1. Suffix-array search: for every (query, offset) the index sa_search
//      reports (lower_bound_kernel), query.c:104-110.
//   2. Streams: per offset the forward suffix at that index and the backward
//      one before it (query.c:112-128).  Every stream is extended on the GPU
//      slot by slot until its first non-seed (stream_round_kernel), in
//      rounds of doubling length shared by all streams still open, each
//      round's elements kept on the device (only each request's first
//      non-seed comes back); the walk pops nearly every element of every
//      stream, so they are computed up front.  compact_kernel then lays each
//      stream's seeds out back to back.
//   3. The walk and hit list of each query on the GPU, one wavefront per
//      query (walk_core.h, walk_kernels.hip); only the hit lists come back.
//      A query whose walk reaches the end of a stream cut at
//      max_stream_elements is searched again with a higher cap.
//   4. Forward Smith-Waterman of every hit whose gapless score is >= 40
//      (ssw_wrapper.c:171-176) through the library's serving loop over the
//      resident proteins; E-values and bitscores on the host with the
//      reference's expressions (ssw_wrapper.c:179-183, 239-245).
//   5. Output order: the hit list's traversal order, stable sort by bitscore,
//      the first -H, then the E-value filter (engine.c:193-234); with -B the
//      traceback statistics of those hits on the GPU (stats_compute).
//
// Schedules: TPZ_SCHED_PERSISTENT puts the whole batch through each stage
// in one set of launches; TPZ_SCHED_PLAIN runs the queries one at a time
// (one set of launches per query), the comparison form.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/topaz_amd.h"
#include "gapless_kernels.h"
#include "runtime.h"
#include "sw_kernels.h"
#include "walk.h"
#include "walk_kernels.h"

namespace tpz {
namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// ---------------------------------------------------------------------------
// grow-only buffers of one search call
// ---------------------------------------------------------------------------
// Device buffers grow in the order of their slot's stream (hipMallocAsync /
// hipFreeAsync): a plain hipFree waits for the whole device, so a search
// growing a buffer would stall until another search's Smith-Waterman launch
// ends.  Growth is by half again the size asked for, so a run of batches of
// similar size settles after a few.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;     // the owning slot's stream
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void* get(size_t n) {
        if (n > bytes) {
            if (p && check(hipFreeAsync(p, st), "hipFreeAsync")) return nullptr;
            p = nullptr;
            bytes = 0;
            const size_t want = std::max(n + n / 2, (size_t)1 << 20);
            if (check(hipMallocAsync(&p, want, st), "hipMallocAsync")) return nullptr;
            bytes = want;
        }
        return p;
    }
};
// Pinned host buffers: no stream-ordered free exists, and hipHostFree also
// waits for the device, so a grown-out buffer is kept until the slot goes.
struct HostBuf {
    char* p = nullptr;
    size_t bytes = 0;
    std::vector<char*> old;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
        for (char* q : old) (void)hipHostFree(q);
    }
    char* get(size_t n) {
        if (n > bytes) {
            if (p) old.push_back(p);
            p = nullptr;
            bytes = 0;
            const size_t want = std::max(n + n / 2, (size_t)1 << 20);
            if (check(hipHostMalloc((void**)&p, want, hipHostMallocDefault), "hipHostMalloc"))
                return nullptr;
            bytes = want;
        }
        return p;
    }
};

// Device and pinned memory of a statistics run (RunMem over grow-only blocks).
struct StatsMem final : RunMem {
    BlockList d, h;
    hipStream_t st = nullptr;
    ~StatsMem() override {
        for (auto& b : d.blocks) (void)hipFree(b.first);
        for (auto& b : h.blocks) (void)hipHostFree(b.first);
    }
    void reset() override { d.reset(); h.reset(); }
    void* dev(size_t bytes) override {
        return d.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            return check(hipMalloc(&p, n), "hipMalloc") ? nullptr : p;
        }, nullptr);
    }
    char* host(size_t bytes) override {
        return h.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            return check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc") ? nullptr : p;
        }, nullptr);
    }
    hipStream_t stream() override { return st; }
    int wait() override { return check(hipStreamSynchronize(st), "stats"); }
};

// Everything one search call uses; calls take one from the database's pool.
struct Slot {
    hipStream_t stream = nullptr;
    DevBuf d_q, d_qoff, d_qlen, d_rq, d_ro, d_idx;
    DevBuf d_qf, d_qr, d_qf_off, d_qr_off;
    // stream rounds: definitions, the round's requests and their ends, and
    // per round its output arrays
    DevBuf d_defs, d_req, d_req_end, d_round_ptrs, d_jobs;
    std::deque<DevBuf> round_vals, round_prot;
    // the walk: contiguous streams, per-query tables, walk state, hits
    DevBuf d_vals, d_prot, d_sbase, d_slen, d_sopen;
    DevBuf d_qstream, d_qcap, d_qmem, d_qout, d_mem;
    DevBuf d_oidx, d_oscore, d_on, d_ostatus, d_ostats;
    HostBuf h_stage, h_req_end, h_oidx, h_oscore;
    StatsMem stats;
    // the stream every device buffer of the slot grows in
    void bind() {
        for (DevBuf* b : {&d_q, &d_qoff, &d_qlen, &d_rq, &d_ro, &d_idx, &d_qf, &d_qr, &d_qf_off,
                          &d_qr_off, &d_defs, &d_req, &d_req_end, &d_round_ptrs, &d_jobs, &d_vals,
                          &d_prot, &d_sbase, &d_slen, &d_sopen, &d_qstream, &d_qcap, &d_qmem,
                          &d_qout, &d_mem, &d_oidx, &d_oscore, &d_on, &d_ostatus, &d_ostats})
            b->st = stream;
    }
    ~Slot() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace
}  // namespace tpz

using namespace tpz;

struct tpz_db {
    uint64_t t_size = 0, n_proteins = 0;
    const char* text = nullptr;          // the caller's (kept alive by the caller)
    const int64_t* sa = nullptr;         // the caller's host suffix array
    std::vector<uint64_t> start;
    std::vector<int32_t> len;
    ProteinIndex proteins;
    tpz_text* dtext = nullptr;           // letter codes, head window, suffix array
    DeviceBuffer d_raw;                  // the text's bytes, for the suffix-array search
    std::vector<int16_t> table;          // gapless scores per (text code, query code)
    DeviceBuffer d_table;
    DeviceBuffer d_pblock, d_pend;       // the protein index on the device (stream_kernel)
    DeviceBuffer d_sa32;                 // the suffix array as uint32 (texts under 4 GiB)
    int64_t sw_part_pairs = 1 << 20;     // batches from this many pairs: stage S in two parts
    tpz_server* sw = nullptr;            // stage S over the proteins
    int8_t matrix[625];
    int32_t gap_open = 11, gap_extend = 1, max_w = 1;
    std::mutex mu;
    std::vector<std::unique_ptr<Slot>> free_slots;
    ~tpz_db() {
        if (sw) tpz_server_destroy(sw);
        if (dtext) tpz_text_destroy(dtext);
    }
};

struct tpz_search_result {
    std::vector<std::vector<tpz_hit>> hits;
    tpz_search_stats stats{};
};

namespace tpz {
namespace {

// lexicographical.h:19-33, 35-74 (aa_table2, upper_aa, upper_aa_only)
struct Maps {
    int8_t aa2[128];
    char upper[256], upper_only[256];
    Maps() {
        static const char internal_order[] = "ACMLJIVTSKRQZEDBNHYFWXP*G";
        static const char matrix_order[] = "ARNDCQEGHILKMFPSTWYVBJZX*";
        for (int c = 0; c < 128; ++c) aa2[c] = 23;
        for (int k = 0; k < 25; ++k) {
            const int8_t idx = (int8_t)(strchr(matrix_order, internal_order[k]) - matrix_order);
            aa2['A' + k] = idx;
            aa2['a' + k] = idx;
        }
        for (int c = 0; c < 256; ++c) {
            const bool keep = (c >= 'A' && c <= 'Y') || c == '*' || c == '\n';
            const bool lower = c >= 'a' && c <= 'y';
            upper[c] = keep ? (char)c : lower ? (char)(c - 32) : 'X';
            upper_only[c] = keep ? (char)c : 'V';
        }
    }
    int8_t idx(unsigned char c) const { return aa2[c & 127]; }
};
const Maps& maps() {
    static const Maps m;
    return m;
}

Slot* take_slot(tpz_db* db) {
    {
        std::lock_guard<std::mutex> lk(db->mu);
        if (!db->free_slots.empty()) {
            Slot* s = db->free_slots.back().release();
            db->free_slots.pop_back();
            return s;
        }
    }
    auto s = std::make_unique<Slot>();
    if (check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate"))
        return nullptr;
    s->stats.st = s->stream;
    s->bind();
    return s.release();
}

void give_slot(tpz_db* db, Slot* s) {
    std::lock_guard<std::mutex> lk(db->mu);
    db->free_slots.emplace_back(s);
}

struct Batch {
    int32_t n = 0;
    const char* q_chars = nullptr;
    const int64_t* q_off = nullptr;
    const int32_t* q_len = nullptr;
};

// Upload the batch's queries in both layouts the kernels read: raw bytes
// (suffix-array search) and doubled letter codes (stream_kernel).
struct QueryData {
    std::vector<int64_t> off;      // offsets into the raw copy
    QueryCodes codes;
};

int upload_queries(tpz_db* db, Slot& sl, const Batch& b, QueryData& qd) {
    int64_t tot = 0;
    qd.off.assign(b.n, 0);
    for (int32_t k = 0; k < b.n; ++k) {
        qd.off[k] = tot;
        tot += b.q_len[k];
    }
    char* h = sl.h_stage.get((size_t)tot + 64);
    if (!h) return -1;
    for (int32_t k = 0; k < b.n; ++k) memcpy(h + qd.off[k], b.q_chars + b.q_off[k], b.q_len[k]);
    void* dq = sl.d_q.get((size_t)tot + 64);
    void* doff = sl.d_qoff.get((size_t)b.n * 8 + 8);
    void* dlen = sl.d_qlen.get((size_t)b.n * 4 + 4);
    if (!dq || !doff || !dlen) return -1;
    if (check(hipMemcpyAsync(dq, h, (size_t)tot, hipMemcpyHostToDevice, sl.stream), "upload") ||
        check(hipMemcpyAsync(doff, qd.off.data(), (size_t)b.n * 8, hipMemcpyHostToDevice,
                             sl.stream), "upload") ||
        check(hipMemcpyAsync(dlen, b.q_len, (size_t)b.n * 4, hipMemcpyHostToDevice, sl.stream),
              "upload"))
        return -1;
    // pageable copies above are staged before the call returns; the codes
    // go the same way
    if (gapless_query_codes(db->dtext->codes, b.n, b.q_chars, b.q_off, b.q_len, qd.codes))
        return -1;
    void* qf = sl.d_qf.get(qd.codes.fwd.size());
    void* qr = sl.d_qr.get(qd.codes.rev.size());
    void* qfo = sl.d_qf_off.get(qd.codes.fwd_off.size() * 8 + 8);
    void* qro = sl.d_qr_off.get(qd.codes.rev_off.size() * 8 + 8);
    if (!qf || !qr || !qfo || !qro) return -1;
    if (check(hipMemcpyAsync(qf, qd.codes.fwd.data(), qd.codes.fwd.size(), hipMemcpyHostToDevice,
                             sl.stream), "upload") ||
        check(hipMemcpyAsync(qr, qd.codes.rev.data(), qd.codes.rev.size(), hipMemcpyHostToDevice,
                             sl.stream), "upload") ||
        check(hipMemcpyAsync(qfo, qd.codes.fwd_off.data(), qd.codes.fwd_off.size() * 8,
                             hipMemcpyHostToDevice, sl.stream), "upload") ||
        check(hipMemcpyAsync(qro, qd.codes.rev_off.data(), qd.codes.rev_off.size() * 8,
                             hipMemcpyHostToDevice, sl.stream), "upload"))
        return -1;
    return check(hipStreamSynchronize(sl.stream), "query upload");
}

GaplessArgs stream_gapless_args(tpz_db* db, Slot& sl, int32_t n_queries, const int32_t* q_len_dev) {
    GaplessArgs a{};
    a.t_fwd = (const uint8_t*)db->dtext->d_fwd.ptr;
    a.t_head = (const uint8_t*)db->dtext->d_head.ptr;
    a.t_size = (int64_t)db->t_size;
    a.sa = db->dtext->has_sa ? (const int64_t*)db->dtext->d_sa.ptr : nullptr;
    a.q_fwd = (const uint8_t*)sl.d_qf.p;
    a.qf_off = (const int64_t*)sl.d_qf_off.p;
    a.q_rev = (const uint8_t*)sl.d_qr.p;
    a.qr_off = (const int64_t*)sl.d_qr_off.p;
    a.q_len = q_len_dev;
    a.n_queries = n_queries;
    a.table = (const int16_t*)db->d_table.ptr;
    return a;
}

struct Options {
    tpz_search_options o;
    int32_t capacity(int32_t qlen) const {
        // hit.c:109-113
        return qlen > 3000 ? (int32_t)(1.5 * o.num_hits) : o.number_of_alignments;
    }
};

// Per-query state through the stages.
struct QueryWork {
    int32_t first_stream = 0, n_streams = 0;   // range in the batch's stream list
    std::vector<uint64_t> hit_index;            // hit list, traversal order
    std::vector<int32_t> hit_gapless;
    bool need = false;                          // walk reached a stream cut at the cap
};

// Upload v into b on the stream (pageable: staged before the call returns).
template <typename T>
T* upload_vec(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    void* p = b.get(v.size() * sizeof(T) + 64);
    if (!p) return nullptr;
    if (!v.empty() &&
        check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st),
              "upload"))
        return nullptr;
    return (T*)p;
}

// The whole search of one batch of queries (all stages).
int search_batch(tpz_db* db, const Options& opt, Slot& sl, const Batch& b,
                 std::vector<std::vector<tpz_hit>>& out, tpz_search_stats& st) {
    const Maps& M = maps();
    const int32_t min_len = opt.o.min_hit_length;
    out.assign(b.n, {});
    double t0 = now_s();
    QueryData qd;
    if (upload_queries(db, sl, b, qd)) return -1;

    // 1. the suffix-array index of every offset (query.c:94-110)
    std::vector<int32_t> rq, ro;
    std::vector<int32_t> rec_begin(b.n + 1, 0);
    for (int32_t k = 0; k < b.n; ++k) {
        rec_begin[k] = (int32_t)rq.size();
        const char* q = b.q_chars + b.q_off[k];
        const int32_t num_suffixes = b.q_len[k] - min_len;
        for (int32_t i = 0; i < num_suffixes; ++i) {
            if (M.upper_only[(unsigned char)q[i]] == 'V') continue;     // _bad_prefix, query.c:86-91
            rq.push_back(k);
            ro.push_back(i);
        }
    }
    rec_begin[b.n] = (int32_t)rq.size();
    const int64_t n_rec = (int64_t)rq.size();
    std::vector<int64_t> idx((size_t)n_rec);
    if (n_rec > 0) {
        void* drq = sl.d_rq.get((size_t)n_rec * 4);
        void* dro = sl.d_ro.get((size_t)n_rec * 4);
        void* didx = sl.d_idx.get((size_t)n_rec * 8);
        if (!drq || !dro || !didx) return -1;
        LowerBoundArgs a{};
        a.text = (const char*)db->d_raw.ptr;
        a.t_size = (int64_t)db->t_size;
        a.sa = db->dtext->has_sa ? (const int64_t*)db->dtext->d_sa.ptr : nullptr;
        a.sa32 = (const uint32_t*)db->d_sa32.ptr;
        a.q_chars = (const char*)sl.d_q.p;
        a.q_off = (const int64_t*)sl.d_qoff.p;
        a.q_len = (const int32_t*)sl.d_qlen.p;
        a.rec_query = (const int32_t*)drq;
        a.rec_offset = (const int32_t*)dro;
        a.n = n_rec;
        a.out = (int64_t*)didx;
        const double g0 = now_s();
        if (check(hipMemcpyAsync(drq, rq.data(), (size_t)n_rec * 4, hipMemcpyHostToDevice,
                                 sl.stream), "upload") ||
            check(hipMemcpyAsync(dro, ro.data(), (size_t)n_rec * 4, hipMemcpyHostToDevice,
                                 sl.stream), "upload") ||
            check(tpz_launch_lower_bound(&a, sl.stream), "lower_bound") ||
            check(hipMemcpyAsync(idx.data(), didx, (size_t)n_rec * 8, hipMemcpyDeviceToHost,
                                 sl.stream), "download") ||
            check(hipStreamSynchronize(sl.stream), "suffix-array search"))
            return -1;
        st.gpu_s += now_s() - g0;
    }
    st.offsets += n_rec;

    // 2. streams in creation order: per offset forward at idx, backward at
    // idx - 1 (a slot outside [0, Tsize) has no suffix)
    std::vector<StreamDef> defs;
    std::vector<QueryWork> qw(b.n);
    for (int32_t k = 0; k < b.n; ++k) {
        qw[k].first_stream = (int32_t)defs.size();
        for (int32_t r = rec_begin[k]; r < rec_begin[k + 1]; ++r) {
            const int64_t pos = idx[r];
            if (pos == -1) continue;
            if (pos < (int64_t)db->t_size) defs.push_back(StreamDef{k, ro[r], pos, 1, 0});
            if (pos != 0 && pos - 1 < (int64_t)db->t_size)
                defs.push_back(StreamDef{k, ro[r], pos - 1, -1, 0});
        }
        qw[k].n_streams = (int32_t)defs.size() - qw[k].first_stream;
    }
    const int64_t n_streams = (int64_t)defs.size();
    if (n_streams >= INT32_MAX) return fail("tpz_search: too many walk streams in one batch");
    st.streams += n_streams;

    // rounds: every open stream gets its next `c` slots, on the device; only
    // each request's first non-seed comes back
    std::vector<int64_t> slen((size_t)n_streams, 0);
    std::vector<uint8_t> sopen((size_t)n_streams, 0);
    struct Piece {
        int32_t stream, round, n;
        int64_t src, at;            // element in the round's output, in the stream
    };
    std::vector<Piece> pieces;
    std::vector<int32_t> open_list((size_t)n_streams);
    for (int64_t s = 0; s < n_streams; ++s) open_list[s] = (int32_t)s;
    const StreamDef* d_defs = upload_vec(sl.d_defs, defs, sl.stream);
    if (!d_defs) return -1;
    const int64_t cap_elems = opt.o.max_stream_elements;
    int64_t have = 0;
    int32_t chunk = std::max(1, opt.o.first_chunk);
    std::vector<const int16_t*> round_vals;
    std::vector<const uint32_t*> round_prot;
    const double g0 = now_s();
    for (int round = 0; !open_list.empty(); ++round) {
        const int64_t n_req = (int64_t)open_list.size();
        const int32_t c = (int32_t)std::min<int64_t>(chunk, cap_elems - have);
        if (n_req >= INT32_MAX) return fail("tpz_search: stream round too large");
        if ((int)sl.round_vals.size() <= round) {
            sl.round_vals.emplace_back();
            sl.round_prot.emplace_back();
            sl.round_vals.back().st = sl.round_prot.back().st = sl.stream;
        }
        int16_t* rv = (int16_t*)sl.round_vals[round].get((size_t)(n_req * c) * 2 + 64);
        uint32_t* rp = (uint32_t*)sl.round_prot[round].get((size_t)(n_req * c) * 4 + 64);
        int32_t* d_end = (int32_t*)sl.d_req_end.get((size_t)n_req * 4 + 64);
        int32_t* h_end = (int32_t*)sl.h_req_end.get((size_t)n_req * 4 + 64);
        const int32_t* d_req = upload_vec(sl.d_req, open_list, sl.stream);
        if (!rv || !rp || !d_end || !h_end || !d_req) return -1;
        round_vals.push_back(rv);
        round_prot.push_back(rp);
        StreamRoundArgs a{};
        a.g = stream_gapless_args(db, sl, b.n, (const int32_t*)sl.d_qlen.p);
        a.defs = d_defs;
        a.req_stream = d_req;
        a.n_req = n_req;
        a.have = have;
        a.chunk = c;
        a.min_len = min_len;
        a.out = rv;
        a.out_protein = rp;
        a.req_end = d_end;
        a.prot_block = (const uint32_t*)db->d_pblock.ptr;
        a.prot_end = (const uint64_t*)db->d_pend.ptr;
        a.sa32 = (const uint32_t*)db->d_sa32.ptr;
        const int64_t waves = n_req * ((c + 63) / 64);
        const int grid = (int)std::max<int64_t>(
            1, std::min<int64_t>((waves + 3) / 4, (int64_t)std::max(num_cu(), 1) * 32));
        if (check(tpz_launch_fill_i32(d_end, c, (int32_t)n_req, sl.stream), "clear ends") ||
            check(tpz_launch_stream_round(&a, grid, sl.stream), "stream_round_kernel") ||
            check(hipMemcpyAsync(h_end, d_end, (size_t)n_req * 4, hipMemcpyDeviceToHost,
                                 sl.stream), "download") ||
            check(hipStreamSynchronize(sl.stream), "stream round"))
            return -1;
        st.elements += n_req * c;
        ++st.rounds;
        std::vector<int32_t> next;
        for (int64_t i = 0; i < n_req; ++i) {
            const int32_t s = open_list[i], e = h_end[i];
            if (e > 0) pieces.push_back(Piece{s, round, e, i * c, have});
            slen[s] += e;
            if (e == c) {
                if (have + c < cap_elems) next.push_back(s);
                else sopen[s] = 1;          // cut at the cap
            }
        }
        open_list.swap(next);
        have += c;
        // each round adds half the slots the open streams hold: a stream that
        // ends inside a round wastes less of it than with doubling; above 64,
        // whole wavefronts of slots
        int64_t next_chunk = std::max<int64_t>(std::max(1, opt.o.first_chunk), have / 2);
        if (next_chunk > 64) next_chunk = (next_chunk + 63) & ~(int64_t)63;
        chunk = (int32_t)std::min<int64_t>(next_chunk, 1 << 20);
    }

    // each stream's seeds back to back, in stream order
    std::vector<int64_t> sbase((size_t)n_streams + 1, 0);
    for (int64_t s = 0; s < n_streams; ++s) sbase[s + 1] = sbase[s] + slen[s];
    const int64_t n_elem = sbase[n_streams];
    std::vector<CompactJob> jobs(pieces.size());
    for (size_t i = 0; i < pieces.size(); ++i) {
        const Piece& pc = pieces[i];
        jobs[i] = CompactJob{pc.src, sbase[pc.stream] + pc.at, pc.n, pc.round};
    }
    std::vector<const void*> ptrs;
    for (auto* v : round_vals) ptrs.push_back(v);
    for (auto* v : round_prot) ptrs.push_back(v);
    int16_t* d_vals = (int16_t*)sl.d_vals.get((size_t)n_elem * 2 + 64);
    uint32_t* d_prot = (uint32_t*)sl.d_prot.get((size_t)n_elem * 4 + 64);
    const CompactJob* d_jobs = upload_vec(sl.d_jobs, jobs, sl.stream);
    const void* const* d_ptrs = upload_vec(sl.d_round_ptrs, ptrs, sl.stream);
    if (!d_vals || !d_prot || !d_jobs || !d_ptrs) return -1;
    {
        CompactArgs a{};
        a.jobs = d_jobs;
        a.n_jobs = (int64_t)jobs.size();
        a.round_vals = (const int16_t* const*)d_ptrs;
        a.round_prot = (const uint32_t* const*)(d_ptrs + round_vals.size());
        a.vals = d_vals;
        a.prot = d_prot;
        if (check(tpz_launch_compact(&a, sl.stream), "compact_kernel")) return -1;
    }
    const double t_streams = now_s();
    st.seed_s += t_streams - t0;

    // 3. the walk of every query, one wavefront each
    const int32_t max_score = 255 * std::max(1, db->max_w);
    std::vector<int32_t> q_stream(b.n + 1), q_cap(b.n);
    std::vector<int64_t> q_mem(b.n + 1, 0), q_out(b.n + 1, 0);
    std::vector<int32_t> slen32((size_t)n_streams);
    for (int64_t s = 0; s < n_streams; ++s) slen32[s] = (int32_t)slen[s];
    for (int32_t k = 0; k < b.n; ++k) {
        q_stream[k] = qw[k].first_stream;
        q_cap[k] = opt.capacity(b.q_len[k]);
        q_mem[k + 1] = q_mem[k] +
                       (int64_t)walk_mem_layout(qw[k].n_streams, q_cap[k], max_score).total;
        q_out[k + 1] = q_out[k] + q_cap[k];
    }
    q_stream[b.n] = (int32_t)n_streams;
    const int64_t n_out = q_out[b.n];
    WalkArgs wa{};
    wa.s_base = upload_vec(sl.d_sbase, sbase, sl.stream);
    wa.s_len = upload_vec(sl.d_slen, slen32, sl.stream);
    wa.s_open = upload_vec(sl.d_sopen, sopen, sl.stream);
    wa.vals = d_vals;
    wa.prot = d_prot;
    wa.q_stream = upload_vec(sl.d_qstream, q_stream, sl.stream);
    wa.q_capacity = upload_vec(sl.d_qcap, q_cap, sl.stream);
    wa.q_mem = upload_vec(sl.d_qmem, q_mem, sl.stream);
    wa.mem = (char*)sl.d_mem.get((size_t)q_mem[b.n] + 256);
    wa.q_out = upload_vec(sl.d_qout, q_out, sl.stream);
    wa.out_index = (uint32_t*)sl.d_oidx.get((size_t)n_out * 4 + 64);
    wa.out_score = (int32_t*)sl.d_oscore.get((size_t)n_out * 4 + 64);
    wa.out_n = (int32_t*)sl.d_on.get((size_t)b.n * 4 + 64);
    wa.out_status = (int32_t*)sl.d_ostatus.get((size_t)b.n * 4 + 64);
    wa.out_stats = (int64_t*)sl.d_ostats.get((size_t)b.n * 24 + 64);
    wa.n_queries = b.n;
    wa.limit = opt.o.number_of_seeds;
    wa.max_score = max_score;
    uint32_t* h_idx = (uint32_t*)sl.h_oidx.get((size_t)n_out * 4 + 64);
    int32_t* h_sc = (int32_t*)sl.h_oscore.get((size_t)n_out * 4 + 64);
    if (!wa.s_base || !wa.s_len || !wa.s_open || !wa.q_stream || !wa.q_capacity || !wa.q_mem ||
        !wa.mem || !wa.q_out || !wa.out_index || !wa.out_score || !wa.out_n || !wa.out_status ||
        !wa.out_stats || !h_idx || !h_sc)
        return -1;
    std::vector<int32_t> h_n(b.n), h_status(b.n);
    std::vector<int64_t> h_stats((size_t)b.n * 3);
    if (check(tpz_launch_walk(&wa, sl.stream), "walk_kernel") ||
        check(hipMemcpyAsync(h_n.data(), wa.out_n, (size_t)b.n * 4, hipMemcpyDeviceToHost,
                             sl.stream), "download") ||
        check(hipMemcpyAsync(h_status.data(), wa.out_status, (size_t)b.n * 4,
                             hipMemcpyDeviceToHost, sl.stream), "download") ||
        check(hipMemcpyAsync(h_stats.data(), wa.out_stats, (size_t)b.n * 24,
                             hipMemcpyDeviceToHost, sl.stream), "download") ||
        check(hipMemcpyAsync(h_idx, wa.out_index, (size_t)n_out * 4, hipMemcpyDeviceToHost,
                             sl.stream), "download") ||
        check(hipMemcpyAsync(h_sc, wa.out_score, (size_t)n_out * 4, hipMemcpyDeviceToHost,
                             sl.stream), "download") ||
        check(hipStreamSynchronize(sl.stream), "walk"))
        return -1;
    st.gpu_s += now_s() - g0;
    std::vector<int32_t> need;
    for (int32_t k = 0; k < b.n; ++k) {
        if (h_status[k] == kWalkBroken)
            return fail("tpz_search: the walk of query %d found its hit list inconsistent", k);
        st.pops += h_stats[3 * (size_t)k];
        st.runs += h_stats[3 * (size_t)k + 1];
        if (h_status[k] == kWalkNeed) {
            qw[k].need = true;
            need.push_back(k);
            continue;
        }
        if (h_n[k] < 0 || h_n[k] > q_cap[k])
            return fail("tpz_search: the walk of query %d returned %d hits", k, h_n[k]);
        st.hits += h_n[k];
    }
    parallel_for_each(b.n, host_threads(), [&](int64_t k) {
        if (qw[k].need) return;
        qw[k].hit_index.assign(h_idx + q_out[k], h_idx + q_out[k] + h_n[k]);
        qw[k].hit_gapless.assign(h_sc + q_out[k], h_sc + q_out[k] + h_n[k]);
    });
    const double t_walk = now_s();
    st.walk_s += t_walk - t_streams;

    // 4. forward Smith-Waterman of the hits with gapless score >= 40
    // queries as matrix rows, and per query its hits with gapless score >= 40
    // as pairs in hit order, on host threads
    std::vector<int64_t> q_off(b.n + 1, 0);
    std::vector<int64_t> hit_base(b.n + 1, 0), pair_base(b.n + 1, 0);
    std::vector<int64_t> q_pairs(b.n, 0);
    parallel_for_each(b.n, host_threads(), [&](int64_t k) {
        const QueryWork& w = qw[k];
        int64_t c = 0;
        for (size_t h = 0; h < w.hit_index.size(); ++h) c += w.hit_gapless[h] >= 40;
        q_pairs[k] = c;
    });
    for (int32_t k = 0; k < b.n; ++k) {
        q_off[k + 1] = q_off[k] + b.q_len[k];
        hit_base[k + 1] = hit_base[k] + (int64_t)qw[k].hit_index.size();
        pair_base[k + 1] = pair_base[k] + q_pairs[k];
    }
    const int64_t n_pairs = pair_base[b.n];
    std::vector<int8_t> q_res((size_t)q_off[b.n]);
    std::vector<int32_t> pq((size_t)n_pairs), ps((size_t)n_pairs);
    std::vector<int64_t> pair_of((size_t)hit_base[b.n]);   // per (query, hit) -> pair number or -1
    parallel_for_each(b.n, host_threads(), [&](int64_t k) {
        const char* q = b.q_chars + b.q_off[k];
        for (int32_t j = 0; j < b.q_len[k]; ++j)                     // ssw_wrapper.c:115-121
            q_res[q_off[k] + j] = M.idx((unsigned char)M.upper[(unsigned char)q[j]]);
        const QueryWork& w = qw[k];
        int64_t p = pair_base[k];
        for (size_t h = 0; h < w.hit_index.size(); ++h) {
            if (w.hit_gapless[h] < 40) {
                pair_of[hit_base[k] + h] = -1;
                continue;
            }
            pair_of[hit_base[k] + h] = p;
            pq[p] = (int32_t)k;
            ps[p] = (int32_t)w.hit_index[h];
            ++p;
        }
    });
    std::vector<int32_t> sc((size_t)n_pairs), re((size_t)n_pairs), rd((size_t)n_pairs);
    // Stage S in parts of whole queries with about equal pair counts,
    // submitted one after the other: the server launches the first part
    // while this thread plans the next, and the output stage of a part
    // (step 5) runs while the parts after it are on the device.  A batch of
    // fewer than sw_part_pairs (1 M) pairs goes in one part: its launches
    // are short, and a second one would add its own tail.
    const int sw_parts = n_pairs >= std::max<int64_t>(1, db->sw_part_pairs) ? 2 : 1;
    std::vector<int32_t> q_cut(sw_parts + 1, b.n);
    q_cut[0] = 0;
    for (int part = 1; part < sw_parts; ++part)
        q_cut[part] = (int32_t)(std::lower_bound(pair_base.begin(), pair_base.end(),
                                                 n_pairs * part / sw_parts) - pair_base.begin());
    for (int part = 1; part <= sw_parts; ++part) q_cut[part] = std::max(q_cut[part], q_cut[part - 1]);
    std::vector<uint64_t> tickets(sw_parts, 0);
    std::vector<uint8_t> submitted(sw_parts, 0);
    auto wait_rest = [&](int from) {           // on an error: collect what was submitted
        for (int part = from; part < sw_parts; ++part) {
            if (!submitted[part]) continue;
            const int64_t p0 = pair_base[q_cut[part]];
            double lat = 0;
            (void)tpz_server_wait(db->sw, tickets[part], sc.data() + p0, re.data() + p0,
                                  rd.data() + p0, &lat);
        }
    };
    for (int part = 0; part < sw_parts; ++part) {
        const int32_t q0 = q_cut[part], q1 = q_cut[part + 1];
        const int64_t p0 = pair_base[q0], np = pair_base[q1] - p0;
        if (np == 0) continue;
        std::vector<int64_t> part_off((size_t)(q1 - q0));
        for (int32_t k = q0; k < q1; ++k) part_off[k - q0] = q_off[k] - q_off[q0];
        std::vector<int32_t> part_q((size_t)np);
        for (int64_t x = 0; x < np; ++x) part_q[x] = pq[p0 + x] - q0;
        if (tpz_server_submit(db->sw, q1 - q0, q_res.data() + q_off[q0], part_off.data(),
                              b.q_len + q0, (int32_t)np, part_q.data(), ps.data() + p0,
                              &tickets[part])) {
            wait_rest(0);
            return -1;
        }
        submitted[part] = 1;
    }
    st.pairs += n_pairs;

    // 5. hits in output order (engine.c:193-234), E-values (ssw_wrapper.c:171-245),
    // in parallel over queries, each part's queries once its stage S is back
    const tpz_search_options& o = opt.o;
    std::vector<std::vector<int32_t>> stat_hits(b.n);   // reported hits needing statistics
    std::vector<int64_t> q_cells(b.n, 0);
    for (int part = 0; part < sw_parts; ++part) {
        const int32_t q0 = q_cut[part], q1 = q_cut[part + 1];
        if (submitted[part]) {
            const int64_t p0 = pair_base[q0];
            double lat = 0;
            submitted[part] = 0;
            if (tpz_server_wait(db->sw, tickets[part], sc.data() + p0, re.data() + p0, rd.data() + p0,
                                &lat)) {
                wait_rest(part + 1);
                return -1;
            }
        }
        parallel_for_each(q1 - q0, host_threads(), [&](int64_t kk) {
            const int32_t k = q0 + (int32_t)kk;
            const QueryWork& w = qw[k];
            const int32_t m = b.q_len[k];
            const double correction = log(o.k * m * (double)o.db_size) / o.H;
            const double corrected_q = fmax(m - correction, 1.0);
            const double corrected_db =
                fmax((double)o.db_size - (correction * (double)o.db_proteins), 1.0);
            // each printable hit's bitscore and E-value; NULL alignments are
            // neither printed nor counted
            const size_t nh = w.hit_index.size();
            thread_local std::vector<int32_t> cand;
            thread_local std::vector<double> bits, ev;
            cand.clear();
            bits.resize(nh);
            ev.resize(nh);
            int64_t cells = 0;
            for (size_t h = 0; h < nh; ++h) {
                const int64_t p = pair_of[hit_base[k] + h];
                if (p < 0) {
                    bits[h] = 0;
                    ev[h] = 1000;
                    cand.push_back((int32_t)h);
                    continue;
                }
                cells += (int64_t)m * db->len[ps[p]];
                if (sc[p] >= 32767) continue;
                const double bs = (o.lambda * sc[p] - log(o.k));
                const double e = corrected_q * corrected_db * exp(-bs);
                ev[h] = e > 1.0e-180 ? e : 0.0;
                bits[h] = bs / M_LN2;
                cand.push_back((int32_t)h);
            }
            q_cells[k] = cells;
            // the first num_hits by bitscore, equal bitscores in hit order: what
            // qsort with compare_bitscore (utils.c:140-152) keeps, since glibc
            // sorts these pointer arrays by merging, which is stable
            const size_t keep = std::min(cand.size(), (size_t)std::max<int64_t>(o.num_hits, 0));
            std::partial_sort(cand.begin(), cand.begin() + keep, cand.end(), [&](int32_t x, int32_t y) {
                return bits[x] > bits[y] || (bits[x] == bits[y] && x < y);
            });
            std::vector<tpz_hit>& res = out[k];
            for (size_t i = 0; i < keep; ++i) {
                const int32_t h = cand[i];
                if (!(ev[h] <= o.evalue)) continue;
                tpz_hit x{};
                x.db_index = w.hit_index[h];
                x.gapless_score = w.hit_gapless[h];
                x.qstart = x.qend = x.sstart = x.send = -1;
                x.length = x.mismatch = x.gapopen = -1;
                const int64_t p = pair_of[hit_base[k] + h];
                if (p < 0) {
                    // not aligned (gapless score < 40): hit_alloc's fields
                    // (hit.c:15-38) but for bitscore and E-value
                    x.rawscore = -1;
                    x.identity = INFINITY;
                } else {
                    x.rawscore = sc[p];
                    x.send = re[p];
                    x.qend = rd[p];
                    x.length = x.gapopen = x.mismatch = 0;
                    x.identity = 0.0;
                }
                x.bitscore = bits[h];
                x.evalue = ev[h];
                res.push_back(x);
            }
            if (!o.superfast)
                for (size_t i = 0; i < res.size(); ++i) stat_hits[k].push_back((int32_t)i);
        });
    }
    const double t_sw = now_s();
    st.align_s += t_sw - t_walk;
    for (int32_t k = 0; k < b.n; ++k) st.cells += q_cells[k];

    // 6. traceback statistics of the reported hits (query.c:250-299)
    if (!o.superfast) {
        std::vector<int32_t> hq, hs, r1, f1, s1;
        std::vector<std::pair<int32_t, int32_t>> who;
        std::vector<int8_t> s_res;
        std::vector<int64_t> s_off;
        std::vector<int32_t> s_len;
        for (int32_t k = 0; k < b.n; ++k)
            for (int32_t i : stat_hits[k]) {
                const tpz_hit& x = out[k][i];
                if (x.rawscore <= 0 || x.rawscore >= 32767) continue;
                const uint64_t pidx = x.db_index;
                s_off.push_back((int64_t)s_res.size());
                s_len.push_back(db->len[pidx]);
                const char* t = db->text + db->start[pidx];
                for (int32_t j = 0; j < db->len[pidx]; ++j) s_res.push_back(M.idx((unsigned char)t[j]));
                hq.push_back(k);
                hs.push_back((int32_t)s_len.size() - 1);
                r1.push_back(x.qend);
                f1.push_back(x.send);
                s1.push_back(x.rawscore);
                who.emplace_back(k, i);
            }
        if (!hq.empty()) {
            StatsInput in{b.n, q_res.data(), q_off.data(), b.q_len, (int32_t)s_len.size(),
                          s_res.data(), s_off.data(), s_len.data(), (int32_t)hq.size(),
                          hq.data(), hs.data(), r1.data(), f1.data(), s1.data(), db->matrix,
                          db->gap_open, db->gap_extend};
            StatsOutput so;
            if (stats_check(in) || stats_compute(in, sl.stats, so)) return -1;
            for (size_t h = 0; h < who.size(); ++h) {
                tpz_hit& x = out[who[h].first][who[h].second];
                x.qstart = so.read_begin[h];
                x.sstart = so.ref_begin[h];
                x.length = so.len[h];
                x.gapopen = so.gapo[h];
                x.mismatch = so.mis[h];
                x.identity = stats_identity(so.ident[h], so.len[h]);
            }
        }
        st.stats_s += now_s() - t_sw;
    }

    // queries whose walk ran into a stream cut at max_stream_elements: the
    // whole search again for them with a cap 16 times higher
    if (!need.empty()) {
        st.refills += (int64_t)need.size();
        Options more = opt;
        more.o.max_stream_elements =
            cap_elems > INT64_MAX / 16 ? INT64_MAX : cap_elems * 16;
        std::vector<int64_t> n_off(need.size());
        std::vector<int32_t> n_len(need.size());
        for (size_t i = 0; i < need.size(); ++i) {
            n_off[i] = b.q_off[need[i]];
            n_len[i] = b.q_len[need[i]];
        }
        std::vector<std::vector<tpz_hit>> again;
        if (search_batch(db, more, sl, Batch{(int32_t)need.size(), b.q_chars, n_off.data(),
                                             n_len.data()}, again, st))
            return -1;
        for (size_t i = 0; i < need.size(); ++i) out[need[i]] = std::move(again[i]);
    }
    return 0;
}

}  // namespace
}  // namespace tpz

extern "C" {

int tpz_db_create(tpz_db** out, const char* text, uint64_t t_size, const int64_t* sa,
                  uint64_t n_proteins, const uint64_t* protein_start, const int32_t* protein_len,
                  const int8_t* matrix, int32_t gap_open, int32_t gap_extend) {
    if (!out) return fail("tpz_db_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    if (!text || !sa || !protein_start || !protein_len || !matrix || t_size == 0)
        return fail("tpz_db_create: null or empty input");
    if (!gaps_valid(gap_open, gap_extend))
        return fail("tpz_db_create: gap penalties must lie in [0, 255] (got %d/%d)", gap_open,
                    gap_extend);
    if (n_proteins > (uint64_t)INT32_MAX) return fail("tpz_db_create: too many proteins");
    auto db = std::make_unique<tpz_db>();
    db->t_size = t_size;
    db->n_proteins = n_proteins;
    db->text = text;
    db->sa = sa;
    db->start.assign(protein_start, protein_start + n_proteins);
    db->len.assign(protein_len, protein_len + n_proteins);
    if (!db->proteins.build(t_size, n_proteins, protein_start, protein_len))
        return fail("tpz_db_create: the proteins do not tile the text (each protein's letters "
                    "and its newline, back to back from position 0)");
    memcpy(db->matrix, matrix, 625);
    db->gap_open = gap_open;
    db->gap_extend = gap_extend;
    for (int i = 0; i < 625; ++i) db->max_w = std::max<int32_t>(db->max_w, matrix[i]);
    // a text under 4 GiB keeps its suffix array as uint32 on the device:
    // half the upload (the largest array) and half the device memory
    // (TOPAZ_AMD_SA64 keeps int64 at any size, the form of texts of 4 GiB and up)
    const bool sa32 = t_size <= 0xffffffffull && !getenv("TOPAZ_AMD_SA64");
    if (tpz_text_create(&db->dtext, text, t_size, sa32 ? nullptr : sa)) return -1;
    if (sa32 && (db->d_sa32.alloc((size_t)t_size * 4) ||
                 stage_to_device(db->d_sa32.ptr, (size_t)t_size * 4,
                                 [&](uint8_t* out, size_t b, size_t e) {
                                     uint32_t* o = reinterpret_cast<uint32_t*>(out);
                                     for (size_t k = b / 4; k < e / 4; ++k) {
                                         const int64_t v = sa[k];
                                         // an entry outside the text: no suffix
                                         o[k - b / 4] = v >= 0 && (uint64_t)v < t_size
                                                            ? (uint32_t)v : 0xffffffffu;
                                     }
                                 })))
        return -1;
    if (db->d_pblock.upload(db->proteins.blocks().data(), db->proteins.blocks().size() * 4) ||
        db->d_pend.upload(db->proteins.ends().data(), db->proteins.ends().size() * 8))
        return -1;
    if (db->d_raw.upload_large(text, t_size)) return -1;
    db->table.assign(kTableRows * kTableCols, 0);
    gapless_table(db->dtext->codes, matrix, db->table.data());
    if (db->d_table.upload(db->table.data(), db->table.size() * 2)) return -1;
    // the proteins as matrix rows (aa_table2 of the text, ssw_wrapper.c:193-201)
    const Maps& M = maps();
    // written whole by the loop below: no zero fill of gigabytes first
    std::unique_ptr<int8_t[]> res(new int8_t[t_size]);
    std::vector<int64_t> off(n_proteins);
    for (uint64_t k = 0; k < n_proteins; ++k) off[k] = (int64_t)protein_start[k];
    parallel_for_each((int64_t)((t_size + (1 << 20) - 1) >> 20), host_threads(), [&](int64_t c) {
        const uint64_t a = (uint64_t)c << 20, e = std::min<uint64_t>(t_size, a + (1 << 20));
        for (uint64_t i = a; i < e; ++i) res[i] = M.idx((unsigned char)text[i]);
    });
    tpz_server_config cfg{};
    cfg.memory_budget = (uint64_t)16 << 30;
    cfg.max_inflight = 2;
    cfg.max_launch_pairs = 64 << 20;
    cfg.waves_per_group = 4;
    if (const char* e = getenv("TOPAZ_AMD_SEARCH_WAVES")) cfg.waves_per_group = atoi(e);
    if (tpz_server_create(&db->sw, (int32_t)n_proteins, res.get(), off.data(), protein_len,
                          matrix, gap_open, gap_extend, &cfg))
        return -1;
    // a launch of under 2 M pairs (a served request alone) is as long as its
    // longest task: 8 wavefronts per staircase halve that; merged batches
    // keep the configured width, which gives the throughput
    if (const char* e = getenv("TOPAZ_AMD_SW_PART_PAIRS")) db->sw_part_pairs = atoll(e);
    int64_t small_pairs = (int64_t)2 << 20;
    if (const char* e = getenv("TOPAZ_AMD_SMALL_LAUNCH_PAIRS")) small_pairs = atoll(e);
    server_small_launches(db->sw, small_pairs);
    if (check(hipStreamSynchronize(stream()), "database upload")) return -1;
    *out = db.release();
    return 0;
}

void tpz_db_destroy(tpz_db* db) {
    delete db;
}

void tpz_search_defaults(tpz_search_options* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->num_hits = 100;                  // options.h:6-23
    o->evalue = 1e-3;
    o->lambda = 0.267;
    o->k = 0.041;
    o->H = 0.140;
    o->min_hit_length = 5;
    o->number_of_alignments = 5000;
    o->number_of_seeds = 300000;
    o->superfast = 1;
    o->first_chunk = 32;
    o->max_stream_elements = 1 << 20;
}

int tpz_search(tpz_db* db, const tpz_search_options* opt, int32_t n_queries, const char* q_chars,
               const int64_t* q_off, const int32_t* q_len, int32_t schedule,
               tpz_search_result** out) {
    if (!db || !opt || !out) return fail("tpz_search: null argument");
    *out = nullptr;
    if (n_queries < 0) return fail("tpz_search: negative query count");
    if (n_queries > 0 && (!q_chars || !q_off || !q_len)) return fail("tpz_search: null queries");
    if (schedule != TPZ_SCHED_PLAIN && schedule != TPZ_SCHED_PERSISTENT)
        return fail("tpz_search: unknown schedule %d", schedule);
    if (opt->min_hit_length < 1 || opt->number_of_alignments < 0 || opt->number_of_seeds < 0 ||
        opt->first_chunk < 1 || opt->max_stream_elements < 1)
        return fail("tpz_search: invalid options");
    for (int32_t k = 0; k < n_queries; ++k)
        if (q_len[k] < 0) return fail("tpz_search: query %d has negative length", k);
    Options o{*opt};
    auto res = std::make_unique<tpz_search_result>();
    res->hits.resize(n_queries);
    Slot* sl = take_slot(db);
    if (!sl) return -1;
    const double t0 = now_s();
    int rc = 0;
    if (schedule == TPZ_SCHED_PERSISTENT) {
        rc = search_batch(db, o, *sl, Batch{n_queries, q_chars, q_off, q_len}, res->hits,
                          res->stats);
    } else {
        for (int32_t k = 0; k < n_queries && !rc; ++k) {
            std::vector<std::vector<tpz_hit>> one;
            rc = search_batch(db, o, *sl, Batch{1, q_chars, q_off + k, q_len + k}, one,
                              res->stats);
            if (!rc) res->hits[k] = std::move(one[0]);
        }
    }
    res->stats.total_s = now_s() - t0;
    give_slot(db, sl);
    if (rc) return -1;
    *out = res.release();
    return 0;
}

int32_t tpz_search_result_count(const tpz_search_result* r, int32_t query) {
    if (!r || query < 0 || query >= (int32_t)r->hits.size()) return -1;
    return (int32_t)r->hits[query].size();
}

const tpz_hit* tpz_search_result_hits(const tpz_search_result* r, int32_t query) {
    if (!r || query < 0 || query >= (int32_t)r->hits.size()) return nullptr;
    return r->hits[query].data();
}

int tpz_search_result_stats(const tpz_search_result* r, tpz_search_stats* st) {
    if (!r || !st) return fail("null argument");
    *st = r->stats;
    return 0;
}

void tpz_search_result_destroy(tpz_search_result* r) { delete r; }

}  // extern "C"

// ---------------------------------------------------------------------------
// serving: continuous batching of search requests over a resident database
// ---------------------------------------------------------------------------
namespace tpz {
namespace {

struct SearchRequest {
    uint64_t id = 0;
    std::vector<char> chars;
    std::vector<int64_t> off;
    std::vector<int32_t> len;
    int64_t offsets = 0;            // query offsets, for the memory estimate
    size_t bytes = 0;               // device bytes admitted for it
    double t_submit = 0, t_done = 0;
    bool done = false;
    int rc = 0;
    std::string err;
    std::unique_ptr<tpz_search_result> res;
};

}  // namespace
}  // namespace tpz

struct tpz_search_server {
    tpz_db* db = nullptr;
    Options opt{};
    tpz_search_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<tpz::SearchRequest*> pending;
    std::unordered_map<uint64_t, tpz::SearchRequest*> reqs;
    size_t in_use = 0, fixed = 0;
    double elements_per_offset = 64;      // running estimate from finished waves
    uint64_t next_id = 1;
    bool stop = false;
    tpz_search_server_stats stats{};
    std::vector<std::thread> lanes;
};

namespace tpz {
namespace {

// Device bytes a request occupies while its wave runs: query copies and
// offset records; per offset two walk streams with their definitions,
// round requests, layout and copy jobs, and their elements (estimated per
// offset from the waves so far) both in the rounds' outputs and laid out
// for the walk; the queries' walk state and hit lists (walk_bytes); and
// their share of the Smith-Waterman request (pairs of at most a full hit
// list per query).
size_t request_device_bytes(const tpz_search_server* s, int64_t residues, int64_t n_queries,
                            int64_t offsets, int64_t walk_bytes) {
    const double per_offset = 250.0 + 12.0 * s->elements_per_offset;
    const int64_t hits = (int64_t)std::max(s->opt.o.number_of_alignments,
                                           (int32_t)(1.5 * s->opt.o.num_hits));
    return (size_t)(8 * residues + per_offset * (double)offsets + (double)walk_bytes +
                    48.0 * hits * n_queries) + 4096;
}

void search_lane(tpz_search_server* s) {
    Slot* sl = take_slot(s->db);
    for (;;) {
        std::vector<SearchRequest*> wave;
        size_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop && s->pending.empty()) {
                    lk.unlock();
                    if (sl) give_slot(s->db, sl);
                    return;
                }
                // admit from the head while the memory and the wave size allow;
                // a lane alone on the device always takes the head
                int64_t queries = 0;
                while (!s->pending.empty()) {
                    SearchRequest* r = s->pending.front();
                    const bool fits = s->in_use + bytes + r->bytes <= s->cfg.memory_budget - s->fixed;
                    const bool room = queries + (int64_t)r->len.size() <= s->cfg.max_wave_queries;
                    if (!wave.empty() && (!fits || !room)) break;
                    if (wave.empty() && !fits && s->in_use > 0) break;
                    wave.push_back(r);
                    bytes += r->bytes;
                    queries += (int64_t)r->len.size();
                    s->pending.pop_front();
                }
                if (!wave.empty()) break;
                s->cv_work.wait(lk);
            }
            s->in_use += bytes;
            s->stats.peak_bytes = std::max<uint64_t>(s->stats.peak_bytes, s->in_use + s->fixed);
        }
        // one merged batch
        std::vector<char> chars;
        std::vector<int64_t> off;
        std::vector<int32_t> len;
        std::vector<size_t> first;
        for (SearchRequest* r : wave) {
            first.push_back(len.size());
            const int64_t base = (int64_t)chars.size();
            chars.insert(chars.end(), r->chars.begin(), r->chars.end());
            for (size_t k = 0; k < r->len.size(); ++k) {
                off.push_back(base + r->off[k]);
                len.push_back(r->len[k]);
            }
        }
        first.push_back(len.size());
        std::vector<std::vector<tpz_hit>> hits;
        tpz_search_stats st{};
        int rc = sl ? search_batch(s->db, s->opt, *sl,
                                   Batch{(int32_t)len.size(), chars.data(), off.data(), len.data()},
                                   hits, st)
                    : -1;
        const std::string err = rc ? std::string(tpz_last_error()) : std::string();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= bytes;
            const double t = now_s() * 1e3;
            int64_t n_off = 0;
            for (size_t w = 0; w < wave.size(); ++w) {
                SearchRequest* r = wave[w];
                n_off += r->offsets;
                r->rc = rc;
                r->err = err;
                if (!rc) {
                    r->res = std::make_unique<tpz_search_result>();
                    for (size_t k = first[w]; k < first[w + 1]; ++k)
                        r->res->hits.push_back(std::move(hits[k]));
                    r->res->stats = st;
                }
                r->t_done = t;
                r->done = true;
            }
            if (!rc && n_off > 0)
                s->elements_per_offset = 0.5 * s->elements_per_offset +
                                         0.5 * (double)st.elements / (double)n_off;
            s->stats.requests_done += wave.size();
            s->stats.waves += 1;
            s->stats.queries_done += len.size();
            s->stats.max_requests_per_wave =
                std::max<uint64_t>(s->stats.max_requests_per_wave, wave.size());
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace
}  // namespace tpz

extern "C" {

int tpz_search_server_create(tpz_search_server** out, tpz_db* db,
                             const tpz_search_options* opt,
                             const tpz_search_server_config* cfg) {
    if (!out || !db || !opt) return fail("tpz_search_server_create: null argument");
    *out = nullptr;
    if (opt->min_hit_length < 1 || opt->number_of_alignments < 0 || opt->number_of_seeds < 0 ||
        opt->first_chunk < 1 || opt->max_stream_elements < 1)
        return fail("tpz_search_server_create: invalid options");
    auto s = std::make_unique<tpz_search_server>();
    s->db = db;
    s->opt = Options{*opt};
    if (cfg) s->cfg = *cfg;
    if (s->cfg.memory_budget == 0) s->cfg.memory_budget = (uint64_t)16 << 30;
    if (s->cfg.lanes <= 0) s->cfg.lanes = 2;
    if (s->cfg.max_wave_queries <= 0) s->cfg.max_wave_queries = 4096;
    // fixed device memory: the Smith-Waterman launch slots' workspace
    tpz_server_stats sws{};
    if (tpz_server_get_stats(db->sw, &sws)) return -1;
    s->fixed = (size_t)sws.workspace_bytes;
    s->stats.fixed_bytes = s->fixed;
    if (s->fixed >= s->cfg.memory_budget)
        return fail("tpz_search_server_create: the budget of %llu bytes does not cover the %zu "
                    "bytes of fixed workspace", (unsigned long long)s->cfg.memory_budget, s->fixed);
    tpz_search_server* raw = s.release();
    for (int k = 0; k < raw->cfg.lanes; ++k) raw->lanes.emplace_back(search_lane, raw);
    *out = raw;
    return 0;
}

int tpz_search_server_submit(tpz_search_server* s, int32_t n_queries, const char* q_chars,
                             const int64_t* q_off, const int32_t* q_len, uint64_t* ticket) {
    if (!s || !ticket) return fail("tpz_search_server_submit: null argument");
    if (n_queries < 0 || (n_queries > 0 && (!q_chars || !q_off || !q_len)))
        return fail("tpz_search_server_submit: bad queries");
    auto r = std::make_unique<SearchRequest>();
    const Maps& M = maps();
    const int32_t max_score = 255 * std::max(1, s->db->max_w);
    int64_t walk_bytes = 0;
    for (int32_t k = 0; k < n_queries; ++k) {
        if (q_len[k] < 0) return fail("query %d has negative length", k);
        r->off.push_back((int64_t)r->chars.size());
        r->len.push_back(q_len[k]);
        r->chars.insert(r->chars.end(), q_chars + q_off[k], q_chars + q_off[k] + q_len[k]);
        int64_t offs = 0;
        for (int32_t i = 0; i + s->opt.o.min_hit_length < q_len[k]; ++i)
            offs += M.upper_only[(unsigned char)q_chars[q_off[k] + i]] != 'V';
        r->offsets += offs;
        const int32_t cap = s->opt.capacity(q_len[k]);
        walk_bytes += (int64_t)walk_mem_layout((int32_t)std::min<int64_t>(2 * offs, INT32_MAX - 1),
                                               cap, max_score).total + 8LL * cap;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    r->bytes = request_device_bytes(s, (int64_t)r->chars.size(), n_queries, r->offsets,
                                    walk_bytes);
    if (r->bytes > s->cfg.memory_budget - s->fixed)
        return fail("tpz_search_server_submit: the request needs about %zu device bytes, more "
                    "than the %llu the budget leaves beside the fixed workspace", r->bytes,
                    (unsigned long long)(s->cfg.memory_budget - s->fixed));
    if (s->stop) return fail("tpz_search_server_submit: server stopped");
    r->id = s->next_id++;
    r->t_submit = now_s() * 1e3;
    *ticket = r->id;
    s->reqs[r->id] = r.get();
    s->pending.push_back(r.release());
    s->cv_work.notify_one();
    return 0;
}

int tpz_search_server_wait(tpz_search_server* s, uint64_t ticket, tpz_search_result** out,
                           double* latency_ms) {
    if (!s || !out) return fail("tpz_search_server_wait: null argument");
    *out = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail("unknown ticket %llu", (unsigned long long)ticket);
    SearchRequest* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->reqs.erase(it);
    lk.unlock();
    std::unique_ptr<SearchRequest> own(r);
    if (r->rc) return fail("%s", r->err.c_str());
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    *out = r->res.release();
    return 0;
}

int tpz_search_server_get_stats(tpz_search_server* s, tpz_search_server_stats* st) {
    if (!s || !st) return fail("null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return 0;
}

void tpz_search_server_destroy(tpz_search_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->lanes) t.join();
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
}

}  // extern "C"
