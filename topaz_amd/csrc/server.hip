// Serving loop of libtopaz_amd: continuous batching of Smith-Waterman
// requests over a device-resident database (tpz_server_* in
// include/topaz_amd.h).
//
// A request's task list (pairs of up to four of its queries per wavefront,
// as the batch path builds it) is planned in the submitting thread.  Requests
// then wait in arrival order.  A scheduler thread owns a few launch slots,
// each with its own stream and grow-only device and pinned host buffers.
// Whenever a slot is free it admits waiting requests from the head of the
// queue while their bytes fit the memory budget, concatenates their task
// lists, orders the work items by cost, uploads everything in one copy, and
// launches the persistent kernel with the results copied back behind it on
// the same stream.  While that launch runs, new
// arrivals queue up and go into the next free slot, whose launch starts on the
// device as the first one's workgroups drain.  A finished slot's results are
// split back to their requests; the latency of a request is the time from its
// submit to the moment its results are on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/topaz_amd.h"
#include "runtime.h"
#include "sw_kernels.h"
#include "sw_lane_kernels.h"

namespace tpz {
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Device allocation that only grows, in the order of its slot's stream
// (hipMallocAsync / hipFreeAsync): a plain hipFree waits for the whole device,
// which would hold a launch back until the other slot's launch ends.  Growth
// is by half again the size asked for.
struct GrowBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;           // the slot's stream
    GrowBuffer() = default;
    GrowBuffer(const GrowBuffer&) = delete;
    GrowBuffer& operator=(const GrowBuffer&) = delete;
    ~GrowBuffer() { if (ptr) (void)hipFree(ptr); }
    int reserve(size_t n) {
        if (n <= bytes) return 0;
        if (ptr) {
            if (check(hipFreeAsync(ptr, st), "hipFreeAsync")) return -1;
            ptr = nullptr;
            bytes = 0;
        }
        const size_t want = std::max(n + n / 2, (size_t)1 << 20);
        if (check(hipMallocAsync(&ptr, want, st), "hipMallocAsync")) return -1;
        bytes = want;
        return 0;
    }
};

// Pinned host allocation that only grows.  hipHostFree also waits for the
// device, so a grown-out buffer is kept until the server goes.
struct PinnedBuffer {
    char* ptr = nullptr;
    size_t bytes = 0;
    std::vector<char*> old;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() {
        if (ptr) (void)hipHostFree(ptr);
        for (char* q : old) (void)hipHostFree(q);
    }
    int reserve(size_t n) {
        if (n <= bytes) return 0;
        if (ptr) old.push_back(ptr);
        ptr = nullptr;
        bytes = 0;
        const size_t want = std::max(n + n / 2, (size_t)1 << 20);
        if (check(hipHostMalloc((void**)&ptr, want, hipHostMallocDefault), "hipHostMalloc"))
            return -1;
        bytes = want;
        return 0;
    }
};

struct Request {
    uint64_t id = 0;
    std::vector<int8_t> q_res;
    std::vector<int64_t> q_off;
    std::vector<int32_t> q_len;
    std::vector<int32_t> pq, ps;
    Tasks tasks;                        // planned at submit
    std::vector<int4> rec;              // the forward kernel's entry records (request-local entries)
    std::vector<int32_t> score, ref, read;
    size_t bytes = 0;
    double t_submit = 0, t_done = 0;
    bool done = false;
    int rc = 0;
    std::string err;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    bool busy = false;
    std::vector<Request*> reqs;
    std::vector<size_t> pair_base;      // first merged pair of each request
    size_t n_pairs = 0;
    size_t bytes = 0;                   // admitted request bytes
    GrowBuffer d_in, d_out, d_lane_ws;
    DeviceBuffer d_ws;
    PinnedBuffer h_in, h_out;
    int error_word = 0;                 // offset (words) of the error flag in d_out
};

// Bytes a request holds on the device while in flight: its queries and pairs,
// its share of the task and item lists, and its results.
size_t request_bytes(size_t q_residues, size_t n_queries, size_t n_pairs) {
    return align16(q_residues) + 12 * n_queries + 72 * n_pairs + 256;
}

}  // namespace
}  // namespace tpz

using namespace tpz;

struct tpz_server {
    int32_t n_subjects = 0, max_subject = 0;
    std::vector<int32_t> s_len;
    std::vector<int2> s_meta;            // per subject: packed offset / 16, length
    int64_t s_pad = 0;                   // offset of the padding after the last subject
    DeviceBuffer d_s_res, d_s_off, d_s_len, d_matrix;
    int32_t gap_open = 11, gap_extend = 1, bias = 0;
    bool lanes = false;                  // gap_open <= gap_extend: the lane kernel
    tpz_server_config cfg{};
    int device = 0;
    // launches of fewer pairs run 8 wavefronts per workgroup, whose
    // staircases halve the longest task's time (a small launch is as long as
    // that); 0: every launch runs cfg.waves_per_group
    int64_t small_launch_pairs = 0;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    size_t mem_in_use = 0;
    uint64_t next_id = 1;
    bool stop = false;
    tpz_server_stats stats{};
    std::thread thread;
};

namespace tpz {
namespace {

// Workgroups of one launch at most with W wavefronts each, and the
// workspace words of one workgroup slot (the lane kernel's grows with the
// launch's longest query, in the slot).
int grid_for(int W) {
    return std::max(num_cu(), 1) *
           std::max(1, std::min(kGroupsPerCu, tpz_sw_forward_groups_per_cu(32, W)));
}
int64_t ws_words_for(const tpz_server* s, int W) {
    return s->lanes ? 0 : (int64_t)W * 2 * (item_columns(s->max_subject) + 4) * 64;
}

// Merge the slot's requests, upload, launch, and queue the result copy.  The
// requests' task lists and entry records were built at submit (in the
// submitting threads); here host threads copy them into the pinned upload
// blob, their query, pair, entry and task numbers moved by each request's
// bases.
int launch_slot(tpz_server* s, Slot& sl) {
    const size_t n_req = sl.reqs.size();
    // bases of each request: queries, residues, pairs (= entries), tasks
    std::vector<size_t> qb(n_req + 1, 0), rb(n_req + 1, 0), pb(n_req + 1, 0), tb(n_req + 1, 0);
    for (size_t k = 0; k < n_req; ++k) {
        const Request* r = sl.reqs[k];
        qb[k + 1] = qb[k] + r->q_len.size();
        rb[k + 1] = rb[k] + r->q_res.size();
        pb[k + 1] = pb[k] + r->pq.size();
        tb[k + 1] = tb[k] + (size_t)r->tasks.count();
    }
    const size_t nq = qb[n_req], nres = rb[n_req], np = pb[n_req], nt = tb[n_req];
    sl.pair_base.assign(pb.begin(), pb.end() - 1);
    sl.n_pairs = np;
    // work items from the merged task costs
    Tasks costs;
    costs.cost.reserve(nt);
    for (const Request* r : sl.reqs) costs.cost.insert(costs.cost.end(), r->tasks.cost.begin(), r->tasks.cost.end());
    std::vector<int32_t> items;
    int32_t n_big = 0;
    const int W = s->small_launch_pairs > 0 && (int64_t)np < s->small_launch_pairs
                      ? 8 : s->cfg.waves_per_group;
    const int grid_w = grid_for(W);
    build_items(costs, grid_w, 1.0, W, items, n_big);
    const int n_items = (int)(items.size() / W);
    const size_t n_rec = s->lanes ? 0 : np;

    // one host blob, one copy: residues | q_off | q_len | pq | ps | index | begin | tquery | items
    // | cols | records
    const size_t o_res = 0;
    const size_t o_qoff = align16(o_res + nres + 16);
    const size_t o_qlen = align16(o_qoff + 8 * nq);
    const size_t o_pq = align16(o_qlen + 4 * nq);
    const size_t o_ps = align16(o_pq + 4 * np);
    const size_t o_index = align16(o_ps + 4 * np);
    const size_t o_begin = align16(o_index + 4 * np);
    const size_t o_tq = align16(o_begin + 4 * (nt + 1));
    const size_t o_items = align16(o_tq + 4 * kTaskQueries * nt);
    const size_t o_cols = align16(o_items + 4 * items.size());
    const size_t o_rec = align16(o_cols + 4 * nt);
    const size_t in_bytes = align16(o_rec + sizeof(int4) * n_rec) + 16;
    // score | ref | read | error | counter | (to 8 bytes) keys
    const size_t o_key_words = (3 * np + 2 + 1) & ~(size_t)1;
    const size_t out_words = 3 * np + 2;
    const size_t dev_out_words = o_key_words + 2 * np;
    if (sl.h_in.reserve(in_bytes) || sl.d_in.reserve(in_bytes) ||
        sl.h_out.reserve(out_words * 4) || sl.d_out.reserve(dev_out_words * 4))
        return -1;
    char* h = sl.h_in.ptr;
    int64_t* h_qoff = (int64_t*)(h + o_qoff);
    int32_t* h_qlen = (int32_t*)(h + o_qlen);
    int32_t* h_pq = (int32_t*)(h + o_pq);
    int32_t* h_ps = (int32_t*)(h + o_ps);
    int32_t* h_index = (int32_t*)(h + o_index);
    int32_t* h_begin = (int32_t*)(h + o_begin);
    int32_t* h_tq = (int32_t*)(h + o_tq);
    int32_t* h_cols = (int32_t*)(h + o_cols);
    int4* h_rec = (int4*)(h + o_rec);
    std::vector<int32_t> q_len(nq);
    for (size_t k = 0; k < n_req; ++k) {
        const Request* r = sl.reqs[k];
        memcpy(h + o_res + rb[k], r->q_res.data(), r->q_res.size());
        for (size_t j = 0; j < r->q_len.size(); ++j) {
            h_qlen[qb[k] + j] = q_len[qb[k] + j] = r->q_len[j];
            h_qoff[qb[k] + j] = (int64_t)rb[k] + r->q_off[j];
        }
        const Tasks& rt = r->tasks;
        for (int32_t j = 0; j < rt.count(); ++j) {
            h_begin[tb[k] + j] = (int32_t)pb[k] + rt.begin[j];
            h_cols[tb[k] + j] = rt.cols[j];
            for (int x = 0; x < kTaskQueries; ++x) {
                const int32_t q = rt.query[(size_t)kTaskQueries * j + x];
                h_tq[(tb[k] + j) * kTaskQueries + x] = q >= 0 ? (int32_t)qb[k] + q : -1;
            }
        }
    }
    h_begin[nt] = (int32_t)np;
    memcpy(h + o_items, items.data(), 4 * items.size());
    // the per-pair arrays, in pieces of 64 K pairs over host threads
    const int64_t n_pieces = (int64_t)((np + 65535) >> 16);
    parallel_for(n_pieces, np >= ((size_t)1 << 20) ? host_threads() : 1, [&](int64_t c0, int64_t c1) {
        size_t k = std::upper_bound(pb.begin(), pb.end(), (size_t)c0 << 16) - pb.begin() - 1;
        for (size_t e = (size_t)c0 << 16, stop = std::min(np, (size_t)c1 << 16); e < stop; ++e) {
            while (e >= pb[k + 1]) ++k;
            const Request* r = sl.reqs[k];
            const size_t x = e - pb[k];
            h_pq[e] = (int32_t)qb[k] + r->pq[x];
            h_ps[e] = r->ps[x];
            h_index[e] = (int32_t)pb[k] + r->tasks.index[x];
            if (n_rec) {
                int4 v = r->rec[x];
                if (v.z >= 0) v.z += (int32_t)pb[k];
                h_rec[e] = v;
            }
        }
    });
    char* d = (char*)sl.d_in.ptr;
    int32_t* dout = (int32_t*)sl.d_out.ptr;
    sl.error_word = (int)(3 * np);
    if (check(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, sl.stream), "upload") ||
        check(tpz_launch_fill_i32(dout + 3 * np, 0, 2, sl.stream), "clear flags") ||
        check(tpz_launch_fill_i32(dout + o_key_words, 0, (int32_t)(2 * np), sl.stream), "clear keys"))
        return -1;
    if (n_items > 0) {
        SwKernelArgs a{};
        a.q_res = (const int8_t*)(d + o_res);
        a.q_off = (const int64_t*)(d + o_qoff);
        a.q_len = (const int32_t*)(d + o_qlen);
        a.s_res = (const int8_t*)s->d_s_res.ptr;
        a.s_pad = s->s_pad;
        a.s_off = (const int64_t*)s->d_s_off.ptr;
        a.s_len = (const int32_t*)s->d_s_len.ptr;
        a.pair_query = (const int32_t*)(d + o_pq);
        a.pair_subject = (const int32_t*)(d + o_ps);
        a.pair_index = (const int32_t*)(d + o_index);
        a.entry_rec = (const int4*)(d + o_rec);
        a.task_cols = (const int32_t*)(d + o_cols);
        a.out_key = (unsigned long long*)(dout + o_key_words);
        a.task_begin = (const int32_t*)(d + o_begin);
        a.task_query = (const int32_t*)(d + o_tq);
        a.item_task = (const int32_t*)(d + o_items);
        a.n_items = n_items;
        a.task_counter = dout + 3 * np + 1;
        a.error = dout + 3 * np;
        a.matrix = (const int8_t*)s->d_matrix.ptr;
        a.gap_open = s->gap_open;
        a.gap_extend = s->gap_extend;
        a.workspace = (uint32_t*)sl.d_ws.ptr;
        a.ws_words_per_slot = ws_words_for(s, W);
        a.out_score = dout;
        a.out_ref = dout + np;
        a.out_read = dout + 2 * np;
        if (s->lanes) {
            // the reference's lanes, every pair of the launch
            int32_t max_q = 0;
            for (int32_t v : q_len) max_q = std::max(max_q, v);
            const int64_t rows = lane_rows(max_q);
            const int blocks = (int)std::min<int64_t>((np + kLaneGroupsPerBlock - 1) / kLaneGroupsPerBlock,
                                                      (int64_t)std::max(num_cu(), 1) * 16);
            if (sl.d_lane_ws.reserve((size_t)blocks * kLaneGroupsPerBlock * 3 * rows * 4 + 64))
                return -1;
            LaneArgs la = lane_args(a, s->bias, (int32_t*)sl.d_lane_ws.ptr, rows);
            la.n_list = (int64_t)np;
            if (check(tpz_launch_sw_lane(&la, blocks, sl.stream), "sw_lane")) return -1;
        } else {
            const int grid = std::min(n_items, grid_w);
            if (check(tpz_launch_sw_forward(&a, grid, 32, W, sl.stream), "sw_forward") ||
                check(tpz_launch_sw_decode(&a, (int32_t)np, sl.stream), "sw_decode"))
                return -1;
        }
    }
    if (check(hipMemcpyAsync(sl.h_out.ptr, dout, out_words * 4, hipMemcpyDeviceToHost, sl.stream),
              "readback") ||
        check(hipEventRecord(sl.ev_done, sl.stream), "event"))
        return -1;
    return 0;
}

// Results of a finished slot back to its requests.  Called with the lock held.
void retire_slot(tpz_server* s, Slot& sl, int launch_rc, const std::string& launch_err) {
    const int32_t* out = (const int32_t*)sl.h_out.ptr;
    const bool timed_out = launch_rc == 0 && sl.n_pairs && out[sl.error_word] != 0;
    const double t = now_ms();
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        Request* r = sl.reqs[k];
        const size_t n = r->pq.size(), b = sl.pair_base[k];
        if (launch_rc) {
            r->rc = -1;
            r->err = launch_err;
        } else if (timed_out) {
            r->rc = -1;
            r->err = "sw_forward: a cross-wavefront wait timed out";
        } else {
            r->score.assign(out + b, out + b + n);
            r->ref.assign(out + sl.n_pairs + b, out + sl.n_pairs + b + n);
            r->read.assign(out + 2 * sl.n_pairs + b, out + 2 * sl.n_pairs + b + n);
        }
        r->t_done = t;
        r->done = true;
        s->stats.requests_done += 1;
        s->stats.pairs_done += n;
    }
    s->mem_in_use -= sl.bytes;
    sl.reqs.clear();
    sl.bytes = 0;
    sl.busy = false;
    s->cv_done.notify_all();
}

void scheduler(tpz_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // retire finished launches
        bool any_busy = false;
        for (auto& p : s->slots) {
            Slot& sl = *p;
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.ev_done);
            if (q == hipSuccess) retire_slot(s, sl, 0, "");
            else if (q != hipErrorNotReady) retire_slot(s, sl, -1, hipGetErrorString(q));
            else any_busy = true;
        }
        if (s->stop && !any_busy) break;
        // admit into a free slot
        Slot* free_slot = nullptr;
        for (auto& p : s->slots)
            if (!p->busy) { free_slot = p.get(); break; }
        if (!s->stop && free_slot && !s->pending.empty()) {
            Slot& sl = *free_slot;
            size_t pairs = 0;
            while (!s->pending.empty()) {
                Request* r = s->pending.front();
                if (s->mem_in_use + r->bytes > s->cfg.memory_budget) break;
                if (!sl.reqs.empty() && pairs + r->pq.size() > (size_t)s->cfg.max_launch_pairs) break;
                s->pending.pop_front();
                sl.reqs.push_back(r);
                sl.bytes += r->bytes;
                s->mem_in_use += r->bytes;
                pairs += r->pq.size();
            }
            if (!sl.reqs.empty()) {
                sl.busy = true;
                s->stats.launches += 1;
                s->stats.max_requests_per_launch =
                    std::max<uint64_t>(s->stats.max_requests_per_launch, sl.reqs.size());
                s->stats.peak_bytes = std::max<uint64_t>(s->stats.peak_bytes, s->mem_in_use);
                lk.unlock();
                const int rc = launch_slot(s, sl);
                const std::string err = rc ? std::string(tpz_last_error()) : std::string();
                lk.lock();
                if (rc) {
                    (void)hipStreamSynchronize(sl.stream);
                    retire_slot(s, sl, rc, err);
                }
                continue;
            }
        }
        if (any_busy)
            s->cv_sched.wait_for(lk, std::chrono::microseconds(50));
        else
            s->cv_sched.wait(lk, [&] { return s->stop || !s->pending.empty(); });
    }
    // requests that never got a launch
    for (Request* r : s->pending) {
        r->rc = -1;
        r->err = "server stopped";
        r->done = true;
    }
    s->pending.clear();
    s->cv_done.notify_all();
}

}  // namespace

void server_small_launches(tpz_server* s, int64_t pairs) {
    if (s) s->small_launch_pairs = pairs;
}

}  // namespace tpz

extern "C" {

int tpz_server_create(tpz_server** out, int32_t n_subjects, const int8_t* s_res,
                      const int64_t* s_off, const int32_t* s_len, const int8_t* matrix,
                      int32_t gap_open, int32_t gap_extend, const tpz_server_config* cfg) {
    if (!out) return fail("tpz_server_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    if (n_subjects < 0) return fail("negative subject count");
    if (!gaps_valid(gap_open, gap_extend))
        return fail("gap penalties must lie in [0, 255] (got %d/%d)", gap_open, gap_extend);
    if (!matrix) return fail("null matrix");
    auto s = std::make_unique<tpz_server>();
    s->n_subjects = n_subjects;
    s->gap_open = gap_open;
    s->gap_extend = gap_extend;
    s->lanes = !textbook_gaps(gap_open, gap_extend);
    for (int i = 0; i < 625; ++i) s->bias = std::max<int32_t>(s->bias, -matrix[i]);
    s->device = device_id();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.memory_budget == 0) s->cfg.memory_budget = (uint64_t)1 << 30;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
    if (s->cfg.max_launch_pairs <= 0) s->cfg.max_launch_pairs = 4000000;
    // 8 wavefronts per staircase: a request's latency is its longest pair's
    if (s->cfg.waves_per_group == 0) s->cfg.waves_per_group = 8;
    if (s->cfg.waves_per_group != 4 && s->cfg.waves_per_group != 8)
        return fail("waves_per_group must be 4 or 8 (got %d)", s->cfg.waves_per_group);
    s->s_len.assign(s_len, s_len + n_subjects);
    // subjects: each starts on a 16-byte boundary, padded with residue 25
    std::vector<int64_t> so(n_subjects);
    int64_t stot = 0;
    for (int32_t k = 0; k < n_subjects; ++k) {
        if (s_len[k] < 0) return fail("subject %d has negative length", k);
        if (s_len[k] > kMaxSwLength) return fail("subject %d is longer than %d residues", k, kMaxSwLength);
        so[k] = stot;
        stot += (s_len[k] + 15) & ~15;
        s->max_subject = std::max(s->max_subject, s_len[k]);
    }
    // the packed residues go to the device as host threads write them (a
    // database's proteins are gigabytes): bytes [b, e) of the packed array
    // start inside subject k = the last with so[k] <= b
    const size_t packed = (size_t)stot + 64;
    std::atomic<int64_t> bad{-1};             // a subject with a residue out of range
    int rc = s->d_s_res.alloc(packed);
    if (!rc)
        rc = stage_to_device(s->d_s_res.ptr, packed, [&](uint8_t* o, size_t b, size_t e) {
            int64_t k = (int64_t)(std::upper_bound(so.begin(), so.end(), (int64_t)b) - so.begin()) - 1;
            for (size_t pos = b; pos < e;) {
                while (k + 1 < n_subjects && (size_t)so[k + 1] <= pos) ++k;
                const size_t stop = k + 1 < n_subjects ? std::min(e, (size_t)so[k + 1]) : e;
                const int64_t j0 = k >= 0 ? (int64_t)pos - so[k] : 0;
                const int64_t len = k >= 0 ? s_len[k] : 0;
                for (int64_t j = j0; pos < stop; ++pos, ++j) {
                    int8_t r = 25;
                    if (j < len) {
                        r = s_res[s_off[k] + j];
                        if (r < 0 || r > 24) { int64_t none = -1; bad.compare_exchange_strong(none, k); r = 25; }
                    }
                    o[pos - b] = (uint8_t)r;
                }
            }
        });
    if (!rc && bad.load() >= 0) return fail("subject %lld has a residue out of range", (long long)bad.load());
    s->s_pad = stot;
    rc |= s->d_s_off.upload(so.data(), so.size() * 8) |
          s->d_s_len.upload(s_len, (size_t)n_subjects * 4) | s->d_matrix.upload(matrix, 625);
    // the forward kernel's workspace, room for a launch of either width
    const size_t ws_bytes = 4 * (size_t)std::max(
        grid_for(s->cfg.waves_per_group) * ws_words_for(s.get(), s->cfg.waves_per_group),
        grid_for(8) * ws_words_for(s.get(), 8));
    for (int k = 0; k < s->cfg.max_inflight && !rc; ++k) {
        auto sl = std::make_unique<Slot>();
        rc |= check(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking), "hipStreamCreate");
        sl->d_in.st = sl->d_out.st = sl->d_lane_ws.st = sl->stream;
        rc |= check(hipEventCreateWithFlags(&sl->ev_done, hipEventDisableTiming), "hipEventCreate");
        rc |= sl->d_ws.alloc(ws_bytes);
        s->stats.workspace_bytes += ws_bytes;
        s->slots.push_back(std::move(sl));
    }
    rc |= check(hipStreamSynchronize(stream()), "upload sync");
    s->s_meta.resize((size_t)n_subjects);
    for (int32_t k = 0; k < n_subjects; ++k)
        s->s_meta[k] = make_int2((int32_t)(uint32_t)(so[k] >> 4), s_len[k]);
    if (rc) {
        for (auto& sl : s->slots) {
            if (sl->ev_done) (void)hipEventDestroy(sl->ev_done);
            if (sl->stream) (void)hipStreamDestroy(sl->stream);
        }
        return -1;
    }
    tpz_server* raw = s.release();
    raw->thread = std::thread(tpz::scheduler, raw);
    *out = raw;
    return 0;
}

int tpz_server_submit(tpz_server* s, int32_t n_queries, const int8_t* q_res,
                      const int64_t* q_off, const int32_t* q_len, int32_t n_pairs,
                      const int32_t* pair_query, const int32_t* pair_subject, uint64_t* ticket) {
    if (!s || !ticket) return fail("null argument");
    if (n_queries < 0 || n_pairs < 0) return fail("negative counts");
    auto r = std::make_unique<Request>();
    size_t tot = 0;
    for (int32_t k = 0; k < n_queries; ++k) {
        if (q_len[k] < 0) return fail("query %d has negative length", k);
        if (q_len[k] > kMaxSwLength) return fail("query %d is longer than %d residues", k, kMaxSwLength);
        tot += (size_t)q_len[k];
    }
    r->q_res.resize(tot);
    r->q_off.resize(n_queries);
    r->q_len.assign(q_len, q_len + n_queries);
    size_t o = 0;
    for (int32_t k = 0; k < n_queries; ++k) {
        for (int32_t j = 0; j < q_len[k]; ++j) {
            const int8_t v = q_res[q_off[k] + j];
            if (v < 0 || v > 24) return fail("query %d residue %d out of range", k, v);
            r->q_res[o + j] = v;
        }
        r->q_off[k] = (int64_t)o;
        o += (size_t)q_len[k];
    }
    // the pairs, on host threads: checked, copied, and each one's subject
    // (offset / 16, length) gathered for the planning
    r->pq.resize((size_t)n_pairs);
    r->ps.resize((size_t)n_pairs);
    std::vector<int2> pair_sub((size_t)n_pairs);
    std::atomic<int64_t> bad{INT64_MAX};          // the lowest pair out of range
    parallel_for(n_pairs, n_pairs >= (1 << 18) ? host_threads() : 1, [&](int64_t a, int64_t b) {
        for (int64_t p = a; p < b; ++p) {
            const int32_t q = pair_query[p], k = pair_subject[p];
            if (q < 0 || q >= n_queries || k < 0 || k >= s->n_subjects) {
                int64_t cur = bad.load();
                while (p < cur && !bad.compare_exchange_weak(cur, p)) {}
                return;
            }
            r->pq[(size_t)p] = q;
            r->ps[(size_t)p] = k;
            pair_sub[(size_t)p] = s->s_meta[(size_t)k];
        }
    });
    if (bad.load() != INT64_MAX) {
        const int64_t p = bad.load();
        return fail("pair %d references query %d / subject %d out of range", (int)p,
                    pair_query[p], pair_subject[p]);
    }
    r->bytes = request_bytes(tot, (size_t)n_queries, (size_t)n_pairs);
    if (r->bytes > s->cfg.memory_budget)
        return fail("request needs %zu device bytes, more than the budget of %llu", r->bytes,
                    (unsigned long long)s->cfg.memory_budget);
    {
        const QueryPairs per = pairs_by_query(r->pq, pair_sub.data(), n_queries);
        build_packed_tasks(per, r->q_len, 32, kTaskQueries, r->tasks);
        if (!s->lanes) build_entry_records(r->tasks, r->pq.data(), pair_sub.data(), r->rec);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail("server stopped");
    r->id = s->next_id++;
    r->t_submit = now_ms();
    *ticket = r->id;
    s->reqs[r->id] = r.get();
    s->pending.push_back(r.release());
    s->cv_sched.notify_one();
    return 0;
}

int tpz_server_wait(tpz_server* s, uint64_t ticket, int32_t* score, int32_t* ref_end,
                    int32_t* read_end, double* latency_ms) {
    if (!s) return fail("null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail("unknown ticket %llu", (unsigned long long)ticket);
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->reqs.erase(it);
    lk.unlock();
    std::unique_ptr<Request> own(r);
    if (r->rc) return fail("%s", r->err.c_str());
    const size_t n = r->score.size();
    if (score && n) memcpy(score, r->score.data(), n * 4);
    if (ref_end && n) memcpy(ref_end, r->ref.data(), n * 4);
    if (read_end && n) memcpy(read_end, r->read.data(), n * 4);
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    return 0;
}

int tpz_server_get_stats(tpz_server* s, tpz_server_stats* st) {
    if (!s || !st) return fail("null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return 0;
}

void tpz_server_destroy(tpz_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_sched.notify_all();
    }
    if (s->thread.joinable()) s->thread.join();
    for (auto& kv : s->reqs) delete kv.second;
    for (auto& sl : s->slots) {
        if (sl->ev_done) (void)hipEventDestroy(sl->ev_done);
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
    }
    delete s;
}

}  // extern "C"
