// Drop-in C ABI of TOPAZ's alignment boundary, backed by the gfx950 kernels.
//
// Exports, with the reference's signatures and struct layouts, the entry
// points its search driver calls on the hot path:
//   ssw.h:87-210           ssw_init, init_destroy, init_destroy2, ssw_align,
//                          align_destroy, ssw_align_stats
//   ssw_wrapper.h:11-14    get_ssw_profile, align_using_ssw_profile,
//                          stats_using_ssw_profile, align_using_ssw
//   search.h:14            sa_hsp
// so that query.c / suffix.c link against this library unchanged.  Each call
// runs on the GPU: the calls of all host threads share the library's stream,
// each thread with its own device and pinned host buffers and its own event
// to wait on; the batch entry points of topaz_amd.h are the fast path for
// new callers.
// Errors are fatal, as in the reference (message + _exit(EXIT_FAILURE)):
// the calls have no error return and must never return altered scores.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/topaz_amd.h"
#include "../../include/topaz_ssw_compat.h"
#include "alphabet.h"
#include "gapless_kernels.h"
#include "runtime.h"
#include "sw_kernels.h"
#include "sw_lane_kernels.h"
#include "traceback_kernels.h"

namespace tpz {
namespace {

[[noreturn]] void die(const char* what) {
    fprintf(stderr, "topaz_amd: %s: %s\n", what, tpz_last_error());
    fflush(stderr);
    _exit(EXIT_FAILURE);
}

void hip_or_die(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        fprintf(stderr, "topaz_amd: %s: %s\n", what, hipGetErrorString(e));
        fflush(stderr);
        _exit(EXIT_FAILURE);
    }
}

constexpr size_t kScratchMin = (size_t)1 << 20;
// Pinned blocks: each hipHostMalloc call costs milliseconds whatever its
// size, and TOPAZ's worker threads all make their first call at once, so
// the first block of a thread comes from a pool pinned 16 blocks at a time.
constexpr size_t kPinnedMin = (size_t)4 << 20;
constexpr int kPinnedPoolBlocks = 16;

std::mutex g_pin_mu;
std::vector<char*> g_pin_free;

char* pooled_pinned_block() {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (g_pin_free.empty()) {
        void* p = nullptr;
        hip_or_die(hipHostMalloc(&p, kPinnedMin * kPinnedPoolBlocks, hipHostMallocCoherent),
                   "hipHostMalloc");
        for (int k = kPinnedPoolBlocks - 1; k >= 0; --k)
            g_pin_free.push_back((char*)p + (size_t)k * kPinnedMin);
    }
    char* b = g_pin_free.back();
    g_pin_free.pop_back();
    return b;
}

// Grow-only buffers owned by one host thread.  A buffer grows to at least
// twice its size, so a run of slowly growing calls reallocates rarely, and
// the buffer it replaces is kept, not freed, until the process ends: hipFree
// and hipHostFree wait for the whole device, which would hold the thread up
// behind every other thread's launches.  The kept buffers add up to less
// than the live one.
struct Scratch {                      // device memory
    void* p = nullptr;
    size_t n = 0;
    std::vector<void*> kept;
    void* get(size_t bytes) {
        if (bytes > n) {
            if (p) kept.push_back(p);
            n = std::max(std::max(bytes, 2 * n), kScratchMin);
            hip_or_die(hipMalloc(&p, n), "hipMalloc");
        }
        return p;
    }
};
// Page-locked, coherent host memory: copies from it need no staging, and
// kernels read their inputs from it and write their results into it
// directly (the event wait after the launch makes the results visible).
struct Pinned {
    void* p = nullptr;
    size_t n = 0;
    std::vector<void*> kept;
    char* get(size_t bytes) {
        if (bytes > n) {
            if (p) kept.push_back(p);
            if (!p && bytes <= kPinnedMin) {
                p = pooled_pinned_block();
                n = kPinnedMin;
            } else {
                n = std::max(std::max(bytes, 2 * n), kPinnedMin);
                hip_or_die(hipHostMalloc(&p, n, hipHostMallocCoherent), "hipHostMalloc");
            }
        }
        return (char*)p;
    }
};

// The streams the calls go to: the library's stream, plus
// TOPAZ_AMD_CALL_STREAMS - 1 more, thread k using stream k mod the count.
// Few streams: a stream per thread would spread the threads over the few
// hardware queues HIP multiplexes streams onto (GPU_MAX_HW_QUEUES, 4 by
// default), where a thread's short launch would wait behind a serving-loop
// launch of milliseconds on the same queue.  With 8 or more hardware
// queues, 2 streams (the default then) spread the threads' enqueues over
// two stream locks without sharing a queue with the serving loop.
hipStream_t call_stream(int ordinal) {
    static std::vector<hipStream_t> streams = [] {
        if (ensure_init()) die("device init");
        std::vector<hipStream_t> v{stream()};
        int n = 1;
        if (const char* q = getenv("GPU_MAX_HW_QUEUES"))
            if (atoi(q) >= 8) n = 2;
        if (const char* e = getenv("TOPAZ_AMD_CALL_STREAMS")) {
            n = atoi(e);
            if (n < 1 || n > 16) {
                fprintf(stderr, "topaz_amd: TOPAZ_AMD_CALL_STREAMS=%s is not 1..16\n", e);
                _exit(EXIT_FAILURE);
            }
        }
        for (int k = 1; k < n; ++k) {
            hipStream_t st;
            hip_or_die(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
            v.push_back(st);
        }
        return v;
    }();
    return streams[(size_t)ordinal % streams.size()];
}

// Per host thread: the buffers of one call, the stream its work goes to and
// the event its wait is on (TOPAZ calls the alignment functions from T
// worker threads at once, engine.c:423-457).
struct ThreadCtx {
    hipEvent_t done = nullptr;
    hipStream_t st = nullptr;
    Scratch in, out, ws;
    Pinned pin;
    // the gapless query copies and score table of the last call, on the
    // device (a query's seeds come in many calls in a row): reused while
    // the text, the query bytes and the matrix are the same
    Scratch g_dev;
    const void* g_text = nullptr;
    std::vector<char> g_query;
    int8_t g_matrix[625] = {0};
    size_t g_qr = 0, g_tab = 0, g_meta = 0;
    ThreadCtx() {
        if (ensure_init()) die("device init");
        static std::atomic<int> threads{0};
        st = call_stream(threads.fetch_add(1));
        // blocking: a waiting thread sleeps instead of spinning on a core
        // that another worker thread's host work could use
        hip_or_die(hipEventCreateWithFlags(&done, hipEventDisableTiming | hipEventBlockingSync),
                   "hipEventCreate");
    }
    // waits until this thread's work so far on its stream has finished
    void finish(const char* what) {
        hip_or_die(hipEventRecord(done, st), what);
        hip_or_die(hipEventSynchronize(done), what);
    }
};

ThreadCtx& ctx() {
    static thread_local ThreadCtx c;
    return c;
}

// Traceback statistics of all host threads go to one worker thread with a
// stream of its own: each turn it merges the requests waiting into one
// stats_compute run, whose launches hold every waiting hit (a wavefront or
// thread per hit, so a hit takes as long among others as alone).  The
// worker's device and pinned blocks are reused from run to run.
struct StatsReq {
    const StatsInput* in;
    StatsOutput* out;
    bool done = false;
    int rc = 0;
    std::string err;
};

struct StatsWorker final : RunMem {
    std::mutex mu;
    std::condition_variable cv_req, cv_done;
    std::deque<StatsReq*> pending;
    bool stop = false;
    std::thread th;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    BlockList d, h;
    // RunMem on the worker's blocks and stream
    void reset() override { d.reset(); h.reset(); }
    void* dev(size_t bytes) override {
        return d.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            hip_or_die(hipMalloc(&p, n), "hipMalloc");
            return p;
        }, nullptr);
    }
    char* host(size_t bytes) override {
        return h.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            hip_or_die(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
            return p;
        }, nullptr);
    }
    hipStream_t stream() override { return st; }
    int wait() override {
        return check(hipEventRecord(ev, st), "statistics") ||
               check(hipEventSynchronize(ev), "statistics");
    }
};

bool same_settings(const StatsInput& a, const StatsInput& b) {
    return a.gap_open == b.gap_open && a.gap_extend == b.gap_extend &&
           memcmp(a.matrix, b.matrix, 625) == 0;
}

// the requests' queries, subjects and hits back to back, numbers shifted
struct Merged {
    std::vector<int8_t> q_res, s_res;
    std::vector<int64_t> q_off, s_off;
    std::vector<int32_t> q_len, s_len, hq, hs, re, fe, sc;
    std::vector<int32_t> first_hit;            // per request, then the total
    StatsInput in{};
};

void merge(const std::vector<StatsReq*>& batch, Merged& m) {
    for (StatsReq* r : batch) {
        const StatsInput& x = *r->in;
        const int32_t q0 = (int32_t)m.q_len.size(), s0 = (int32_t)m.s_len.size();
        int64_t qtot = 0, stot = 0;
        for (int32_t k = 0; k < x.n_queries; ++k) qtot = std::max<int64_t>(qtot, x.q_off[k] + x.q_len[k]);
        for (int32_t k = 0; k < x.n_subjects; ++k) stot = std::max<int64_t>(stot, x.s_off[k] + x.s_len[k]);
        const int64_t qb = (int64_t)m.q_res.size(), sb = (int64_t)m.s_res.size();
        m.q_res.insert(m.q_res.end(), x.q_res, x.q_res + qtot);
        m.s_res.insert(m.s_res.end(), x.s_res, x.s_res + stot);
        for (int32_t k = 0; k < x.n_queries; ++k) {
            m.q_off.push_back(qb + x.q_off[k]);
            m.q_len.push_back(x.q_len[k]);
        }
        for (int32_t k = 0; k < x.n_subjects; ++k) {
            m.s_off.push_back(sb + x.s_off[k]);
            m.s_len.push_back(x.s_len[k]);
        }
        m.first_hit.push_back((int32_t)m.hq.size());
        for (int32_t h = 0; h < x.n_hits; ++h) {
            m.hq.push_back(q0 + x.hit_query[h]);
            m.hs.push_back(s0 + x.hit_subject[h]);
            m.re.push_back(x.read_end1[h]);
            m.fe.push_back(x.ref_end1[h]);
            m.sc.push_back(x.score1[h]);
        }
    }
    m.first_hit.push_back((int32_t)m.hq.size());
    const StatsInput& x = *batch[0]->in;
    m.in = StatsInput{(int32_t)m.q_len.size(), m.q_res.data(), m.q_off.data(), m.q_len.data(),
                      (int32_t)m.s_len.size(), m.s_res.data(), m.s_off.data(), m.s_len.data(),
                      (int32_t)m.hq.size(), m.hq.data(), m.hs.data(), m.re.data(), m.fe.data(),
                      m.sc.data(), x.matrix, x.gap_open, x.gap_extend};
}

void stats_loop(StatsWorker* w) {
    (void)hipSetDevice(device_id());
    std::unique_lock<std::mutex> lk(w->mu);
    for (;;) {
        w->cv_req.wait(lk, [&] { return w->stop || !w->pending.empty(); });
        if (w->pending.empty()) break;                 // stopping
        // the waiting requests with the first one's matrix and gaps
        std::vector<StatsReq*> batch;
        for (auto it = w->pending.begin(); it != w->pending.end();) {
            if (batch.empty() || same_settings(*(*it)->in, *batch[0]->in)) {
                batch.push_back(*it);
                it = w->pending.erase(it);
            } else {
                ++it;
            }
        }
        lk.unlock();
        Merged m;
        merge(batch, m);
        StatsOutput o;
        const int rc = stats_compute(m.in, *w, o);
        const std::string err = rc ? std::string(tpz_last_error()) : std::string();
        for (size_t b = 0; b < batch.size() && !rc; ++b) {
            const int32_t h0 = m.first_hit[b], h1 = m.first_hit[b + 1];
            StatsOutput& out = *batch[b]->out;
            out.rc.assign(o.rc.begin() + h0, o.rc.begin() + h1);
            out.read_begin.assign(o.read_begin.begin() + h0, o.read_begin.begin() + h1);
            out.ref_begin.assign(o.ref_begin.begin() + h0, o.ref_begin.begin() + h1);
            out.len.assign(o.len.begin() + h0, o.len.begin() + h1);
            out.gapo.assign(o.gapo.begin() + h0, o.gapo.begin() + h1);
            out.mis.assign(o.mis.begin() + h0, o.mis.begin() + h1);
            out.ident.assign(o.ident.begin() + h0, o.ident.begin() + h1);
        }
        lk.lock();
        for (StatsReq* r : batch) {
            r->rc = rc;
            r->err = err;
            r->done = true;
        }
        w->cv_done.notify_all();
    }
}

StatsWorker* g_stats = nullptr;
std::mutex g_stats_mu;

// Stops the worker before the HIP runtime is torn down at exit.
void stop_stats_worker() {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    if (!g_stats) return;
    {
        std::lock_guard<std::mutex> wl(g_stats->mu);
        g_stats->stop = true;
    }
    g_stats->cv_req.notify_all();
    if (g_stats->th.joinable()) g_stats->th.join();
}

StatsWorker& stats_worker() {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    if (!g_stats) {
        if (ensure_init()) die("device init");
        auto* w = new StatsWorker();
        hip_or_die(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking), "hipStreamCreate");
        hip_or_die(hipEventCreateWithFlags(&w->ev, hipEventDisableTiming | hipEventBlockingSync),
                   "hipEventCreate");
        w->th = std::thread(stats_loop, w);
        g_stats = w;
        atexit(stop_stats_worker);
    }
    return *g_stats;
}

// Traceback statistics of hits through the worker (dies on error).
void stats_run(const StatsInput& in, StatsOutput& out, const char* who) {
    if (stats_check(in)) die(who);
    if (in.n_hits == 0) {
        out = StatsOutput{};
        return;
    }
    StatsWorker& w = stats_worker();
    StatsReq r;
    r.in = &in;
    r.out = &out;
    {
        std::unique_lock<std::mutex> lk(w.mu);
        w.pending.push_back(&r);
        w.cv_req.notify_one();
        w.cv_done.wait(lk, [&] { return r.done; });
    }
    if (r.rc) {
        fprintf(stderr, "topaz_amd: %s: %s\n", who, r.err.c_str());
        fflush(stderr);
        _exit(EXIT_FAILURE);
    }
}

int8_t matrix_bias(const int8_t* mat, int32_t n) {
    int32_t b = 0;                                   // ssw.c:805-811
    for (int32_t i = 0; i < n * n; ++i) b = std::min<int32_t>(b, mat[i]);
    return (int8_t)(-b);
}

// Residues are matrix indices 0..24 (aa_table2 never gives anything else);
// the reference would index its profile out of bounds with another value.
[[noreturn]] void residue_out_of_range(const char* what, int32_t i, int v) {
    fprintf(stderr, "topaz_amd: ssw_align: %s residue %d is %d (not a matrix index 0..24)\n",
            what, i, v);
    fflush(stderr);
    _exit(EXIT_FAILURE);
}

// Forward pass of one pair on the GPU: (score capped at 32767, ref_end1,
// read_end1) as the byte/word dispatch of ssw.c:861-876 would give.
void forward_one(const int8_t* read, int32_t m, const int8_t* ref, int32_t n,
                 const int8_t* mat, int go, int ge, int32_t out[3]) {
    ThreadCtx& c = ctx();
    // the subject starts on a 16-byte boundary and is padded with residue 25
    const size_t m_pad = (((size_t)m + 15) & ~(size_t)15) + 16, n_pad = (((size_t)n + 15) & ~(size_t)15) + 16;
    // layout of the input block
    const size_t o_q = 0, o_s = o_q + m_pad, o_mat = o_s + n_pad, o_meta = (o_mat + 625 + 15) & ~(size_t)15;
    const size_t bytes = o_meta + 128;
    const size_t o_out = (bytes + 15) & ~(size_t)15;
    char* h = c.pin.get(o_out + 16);
    memset(h, 0, bytes);
    memcpy(h + o_q, read, (size_t)m);
    memset(h + o_s, 25, n_pad);
    if (m > kMaxSwLength || n > kMaxSwLength) {
        fprintf(stderr, "topaz_amd: sequences of more than %d residues are not supported\n",
                kMaxSwLength);
        fflush(stderr);
        _exit(EXIT_FAILURE);
    }
    for (int32_t i = 0; i < m; ++i)
        if (read[i] < 0 || read[i] > 24) residue_out_of_range("query", i, read[i]);
    for (int32_t i = 0; i < n; ++i)
        if (ref[i] < 0 || ref[i] > 24) residue_out_of_range("subject", i, ref[i]);
    memcpy(h + o_s, ref, (size_t)n);
    memcpy(h + o_mat, mat, 625);
    int64_t* meta64 = reinterpret_cast<int64_t*>(h + o_meta);
    meta64[0] = 0;                 // q_off
    meta64[1] = 0;                 // s_off
    int32_t* meta32 = reinterpret_cast<int32_t*>(h + o_meta + 16);
    meta32[0] = m;                 // q_len
    meta32[1] = n;                 // s_len
    meta32[2] = 0;                 // pair_query
    meta32[3] = 0;                 // pair_subject
    meta32[4] = 0;                 // task_begin[0]
    meta32[5] = 1;                 // task_begin[1]
    meta32[6] = 0;                 // item 0: task 0 on wavefront 0
    meta32[7] = -1;
    meta32[8] = -1;
    meta32[9] = -1;
    meta32[10] = 0;                // error flag
    meta32[11] = 0;                // task_query of task 0: query 0 only
    meta32[12] = -1;
    meta32[13] = -1;
    meta32[14] = -1;
    meta32[15] = item_columns(n);  // task_cols of task 0
    int32_t* rec32 = reinterpret_cast<int32_t*>(h + o_meta + 80);
    rec32[0] = 0;                  // entry 0: the subject at s_res + 0
    rec32[1] = n;
    rec32[2] = -1;                 // no further pair in its run
    rec32[3] = 0;                  // query slot 0
    char* d = (char*)c.in.get(bytes);
    int32_t* dout = (int32_t*)c.out.get(64);
    // four wavefront buffers (textbook gaps) or the lane kernel's three rows
    const bool lanes = !textbook_gaps(go, ge);
    const size_t ws_words = lanes ? (size_t)3 * lane_rows(m) + 16 : (size_t)4 * 2 * ((size_t)item_columns(n) + 4) * 64;
    uint32_t* dws = (uint32_t*)c.ws.get(ws_words * 4);
    hip_or_die(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.st), "upload");
    SwKernelArgs a{};
    a.q_res = (const int8_t*)(d + o_q);
    a.q_off = (const int64_t*)(d + o_meta);
    a.q_len = (const int32_t*)(d + o_meta + 16);
    a.s_res = (const int8_t*)(d + o_s);
    a.s_off = (const int64_t*)(d + o_meta + 8);
    a.s_len = (const int32_t*)(d + o_meta + 20);
    a.pair_query = (const int32_t*)(d + o_meta + 24);
    a.pair_subject = (const int32_t*)(d + o_meta + 28);
    a.pair_index = nullptr;
    a.entry_rec = (const int4*)(d + o_meta + 80);
    a.task_cols = (const int32_t*)(d + o_meta + 60 + 16);
    a.out_key = (unsigned long long*)(dout + 4);
    a.s_pad = (int64_t)(n_pad - 16);   // 16 residues 25 after the subject
    a.task_begin = (const int32_t*)(d + o_meta + 32);
    a.task_query = (const int32_t*)(d + o_meta + 60);
    a.item_task = (const int32_t*)(d + o_meta + 40);
    a.n_items = 1;
    a.task_counter = nullptr;
    a.error = (int32_t*)(d + o_meta + 56);
    a.matrix = (const int8_t*)(d + o_mat);
    a.gap_open = go;
    a.gap_extend = ge;
    a.workspace = dws;
    a.ws_words_per_slot = (int64_t)ws_words;
    a.out_score = dout;
    a.out_ref = dout + 1;
    a.out_read = dout + 2;
    if (lanes) {
        LaneArgs la = lane_args(a, matrix_bias(mat, 25), (int32_t*)dws, lane_rows(m));
        la.n_list = 1;
        hip_or_die(tpz_launch_sw_lane(&la, 1, c.st), "sw_lane");
    } else {
        hip_or_die(tpz_launch_fill_i32(dout + 4, 0, 2, c.st), "clear key");
        hip_or_die(tpz_launch_sw_forward(&a, 1, 32, 4, c.st), "sw_forward");
        hip_or_die(tpz_launch_sw_decode(&a, 1, c.st), "sw_decode");
    }
    int32_t* ho = (int32_t*)(h + o_out);
    hip_or_die(hipMemcpyAsync(ho, dout, 12, hipMemcpyDeviceToHost, c.st), "readback");
    c.finish("sw_forward");
    memcpy(out, ho, 12);
}

void check_gaps(int go, int ge) {
    if (!gaps_valid(go, ge)) {
        fprintf(stderr, "topaz_amd: gap penalties %d/%d out of range (the reference takes "
                        "them as uint8_t, 0..255)\n", go, ge);
        _exit(EXIT_FAILURE);
    }
}

}  // namespace
}  // namespace tpz

using namespace tpz;

extern "C" {

// ---------------------------------------------------------------------------
// ssw.h
// ---------------------------------------------------------------------------
s_profile* ssw_init(const int8_t* read, const int32_t readLen, const int8_t* mat,
                    const int32_t n, const int8_t score_size) {
    if (n != 25) {
        fprintf(stderr, "topaz_amd: ssw_init: only 25 x 25 matrices are supported\n");
        _exit(EXIT_FAILURE);
    }
    if (score_size != 2) {
        // TOPAZ always asks for both profiles (ssw_wrapper.c:123-127)
        fprintf(stderr, "topaz_amd: ssw_init: only score_size 2 is supported\n");
        _exit(EXIT_FAILURE);
    }
    s_profile* p = (s_profile*)calloc(1, sizeof(s_profile));
    p->read = (int8_t*)read;
    p->mat = mat;
    p->readLen = readLen;
    p->n = n;
    p->bias = (uint8_t)matrix_bias(mat, n);
    // the reference's profile arrays are built on the device per call;
    // non-null markers keep callers that test them working
    p->profile_byte = (void*)p;
    p->profile_word = (void*)p;
    return p;
}

void init_destroy(s_profile* p) { free(p); }

void init_destroy2(s_profile* p) {
    if (!p) return;
    free(p->read);
    free(p);
}

s_align* ssw_align(const s_profile* prof, const int8_t* ref, int32_t refLen,
                   const uint8_t weight_gapO, const uint8_t weight_gapE, const uint8_t flag,
                   const uint16_t filters, const int32_t filterd, const int32_t maskLen) {
    (void)flag; (void)filters; (void)filterd;
    check_gaps(weight_gapO, weight_gapE);
    int32_t r[3];
    forward_one(prof->read, prof->readLen, ref, refLen, prof->mat, weight_gapO, weight_gapE, r);
    if (r[0] >= 32767) return NULL;                  // ssw.c:866-870
    s_align* a = (s_align*)calloc(1, sizeof(s_align));
    a->ref_begin1 = -1;
    a->read_begin1 = -1;
    a->score1 = (uint16_t)r[0];
    a->ref_end1 = r[1];
    a->read_end1 = r[2];
    // the suboptimal search is disabled in the reference (ssw.c:309-328):
    // bests[1] stays zero
    a->score2 = 0;
    a->ref_end2 = maskLen >= 15 ? 0 : -1;
    return a;
}

void align_destroy(s_align* a) {
    if (!a) return;
    free(a->cigar);
    free(a);
}

int ssw_align_stats(const s_profile* prof, const int8_t* ref, int32_t refLen,
                    const uint8_t weight_gapO, const uint8_t weight_gapE, const int32_t maskLen,
                    int32_t read_end1, int32_t ref_end1, uint16_t score1, int32_t* read_begin1,
                    int32_t* ref_begin1, int32_t* length, int32_t* gapopen, int32_t* mismatch,
                    double* identity) {
    (void)maskLen;
    check_gaps(weight_gapO, weight_gapE);
    const int64_t qo = 0, so = 0;
    const int32_t ql = prof->readLen, sl = refLen, zero = 0;
    const int32_t re = read_end1, fe = ref_end1, sc = score1;
    StatsOutput o;
    stats_run(StatsInput{1, prof->read, &qo, &ql, 1, ref, &so, &sl, 1, &zero, &zero, &re, &fe, &sc,
                         prof->mat, weight_gapO, weight_gapE}, o, "ssw_align_stats");
    if (read_begin1) *read_begin1 = o.read_begin[0];
    if (ref_begin1) *ref_begin1 = o.ref_begin[0];
    if (length) *length = o.len[0];
    if (gapopen) *gapopen = o.gapo[0];
    if (mismatch) *mismatch = o.mis[0];
    if (identity) *identity = stats_identity(o.ident[0], o.len[0]);
    return o.rc[0];
}

// ---------------------------------------------------------------------------
// ssw_wrapper.h
// ---------------------------------------------------------------------------
s_profile* get_ssw_profile(seq_t* s1, options_t* opt) {
    const Alphabet& ab = alphabet();
    int8_t* num = (int8_t*)malloc((size_t)s1->seq_len + 1);
    for (int i = 0; i < s1->seq_len; ++i)                 // ssw_wrapper.c:115-121
        num[i] = ab.idx(ab.upper[(unsigned char)s1->seq_buf.buf[i]]);
    return ssw_init(num, s1->seq_len, *opt->substitution, 25, 2);
}

static void karlin(hit_t* hit, const options_t* opt, int32_t qlen, double* correction,
                   double* corrected_q, double* corrected_db) {
    (void)hit;
    *correction = log(opt->k * qlen * opt->db_size) / opt->H;       // ssw_wrapper.c:179-181
    *corrected_q = fmax(qlen - *correction, 1.0);
    *corrected_db = fmax(opt->db_size - (*correction * opt->db_proteins), 1.0);
}

void align_using_ssw_profile(s_profile* profile, seq_t* s2, hit_t* hit, options_t* opt) {
    if (hit->gapless_score < 40) {                                  // ssw_wrapper.c:171-176
        hit->bitscore = 0;
        hit->evalue = 1000;
        hit->is_aligned = 1;
        return;
    }
    double correction, corrected_q, corrected_db;
    karlin(hit, opt, profile->readLen, &correction, &corrected_q, &corrected_db);
    const Alphabet& ab = alphabet();
    std::vector<int8_t> ref((size_t)s2->seq_len + 1);
    for (int i = 0; i < s2->seq_len; ++i) ref[i] = ab.idx(s2->seq_buf.buf[i]);
    s_align* result = ssw_align(profile, ref.data(), s2->seq_len, (uint8_t)opt->gap_open,
                                (uint8_t)opt->gap_extend, 0, 0, 0, 15);
    if (!result) {                                                  // ssw_wrapper.c:222-226
        hit->rawscore = 32767;
        return;
    }
    hit->sstart = result->ref_begin1;
    hit->send = result->ref_end1;
    hit->qstart = result->read_begin1;
    hit->qend = result->read_end1;
    hit->rawscore = result->score1;
    hit->length = 0;
    hit->gapopen = 0;
    hit->mismatch = 0;
    hit->identity = 0.0;
    hit->bitscore = (opt->lambda * hit->rawscore - log(opt->k));     // ssw_wrapper.c:239-245
    hit->evalue = corrected_q * corrected_db * exp(-hit->bitscore);
    hit->evalue = hit->evalue > 1.0e-180 ? hit->evalue : 0.0;
    hit->bitscore /= M_LN2;
    hit->is_aligned = 1;
    hit->effective_length = profile->readLen - correction;
    align_destroy(result);
}

void stats_using_ssw_profile(s_profile* profile, seq_t* s2, hit_t* hit, options_t* opt) {
    if (hit->rawscore == 32767) return;                             // ssw_wrapper.c:136-138
    const Alphabet& ab = alphabet();
    std::vector<int8_t> ref((size_t)s2->seq_len + 1);
    for (int i = 0; i < s2->seq_len; ++i) ref[i] = ab.idx(s2->seq_buf.buf[i]);
    ssw_align_stats(profile, ref.data(), s2->seq_len, (uint8_t)opt->gap_open,
                    (uint8_t)opt->gap_extend, 15, hit->qend, hit->send, (uint16_t)hit->rawscore,
                    &hit->qstart, &hit->sstart, &hit->length, &hit->gapopen, &hit->mismatch,
                    &hit->identity);
}

void align_using_ssw(seq_t* s1, seq_t* s2, hit_t* hit, options_t* opt) {
    double correction, corrected_q, corrected_db;
    karlin(hit, opt, s1->seq_len, &correction, &corrected_q, &corrected_db);
    const Alphabet& ab = alphabet();
    std::vector<int8_t> num((size_t)s1->seq_len + 1), ref((size_t)s2->seq_len + 1);
    for (int i = 0; i < s1->seq_len; ++i) num[i] = ab.idx(ab.upper[(unsigned char)s1->seq_buf.buf[i]]);
    for (int i = 0; i < s2->seq_len; ++i) ref[i] = ab.idx(ab.upper[(unsigned char)s2->seq_buf.buf[i]]);
    s_profile* profile = ssw_init(num.data(), s1->seq_len, *opt->substitution, 25, 2);
    s_align* result = ssw_align(profile, ref.data(), s2->seq_len, (uint8_t)opt->gap_open,
                                (uint8_t)opt->gap_extend, 1, 0, 0, 15);
    if (!result) {
        // the reference dereferences the NULL result here (ssw_wrapper.c:373)
        fprintf(stderr, "topaz_amd: align_using_ssw: alignment score overflow\n");
        _exit(EXIT_FAILURE);
    }
    hit->sstart = result->ref_begin1 + 1;                           // ssw_wrapper.c:373-429
    hit->send = result->ref_end1 + 1;
    hit->qstart = result->read_begin1 + 1;
    hit->qend = result->read_end1 + 1;
    hit->length = 0;
    hit->gapopen = 0;
    hit->mismatch = 0;
    // no CIGAR is produced, so the identity loop runs zero times and the
    // reference divides 0 by (float) 0
    volatile float zero = 0.0f;
    hit->identity = 0 / zero;
    hit->rawscore = result->score1;
    hit->bitscore = (opt->lambda * hit->rawscore - log(opt->k));
    hit->evalue = corrected_q * corrected_db * exp(-hit->bitscore);
    hit->evalue = hit->evalue > 1.0e-180 ? hit->evalue : 0.0;
    hit->bitscore /= M_LN2;
    hit->is_aligned = 1;
    hit->effective_length = s1->seq_len - correction;
    align_destroy(result);
    init_destroy(profile);
}

// ---------------------------------------------------------------------------
// search.h: sa_hsp on a device-resident copy of db->T / db->SA
// ---------------------------------------------------------------------------
namespace {
// db->T (and db->SA) uploaded once per db_t and kept on the device as the
// gapless kernel's text.
struct ResidentText {
    const char* host_T = nullptr;
    uint64_t t_size = 0;
    const int64_t* host_SA = nullptr;
    tpz_text* text = nullptr;
    char* raw = nullptr;            // db->T's bytes, for the suffix-array search
};
std::mutex g_text_mu;
std::vector<ResidentText*> g_texts;

static const ResidentText& resident_text(const db_t* db) {
    std::lock_guard<std::mutex> lk(g_text_mu);
    for (auto* t : g_texts)
        if (t->host_T == db->T && t->t_size == db->Tsize && t->host_SA == db->SA) return *t;
    auto* r = new ResidentText();
    r->host_T = db->T;
    r->t_size = db->Tsize;
    r->host_SA = db->SA;
    if (tpz_text_create(&r->text, db->T, db->Tsize, db->SA)) {
        fprintf(stderr, "topaz_amd: sa_hsp: %s\n", tpz_last_error());
        _exit(EXIT_FAILURE);
    }
    g_texts.push_back(r);
    return *r;
}

// db->T's bytes on the device, uploaded at the first suffix-array search
const char* resident_raw(const db_t* db) {
    const ResidentText& rt = resident_text(db);
    std::lock_guard<std::mutex> lk(g_text_mu);
    ResidentText& w = const_cast<ResidentText&>(rt);
    if (!w.raw) {
        char* d = nullptr;
        hip_or_die(hipMalloc(&d, (size_t)db->Tsize + 1), "hipMalloc");
        hip_or_die(hipMemcpy(d, db->T, (size_t)db->Tsize, hipMemcpyHostToDevice), "text upload");
        w.raw = d;
    }
    return w.raw;
}
}  // namespace

namespace {
// n seeds of one query against the resident text: one launch that reads the
// seeds from pinned host memory and writes the results there (and one upload
// of the query copies and score table when the query or matrix changed).
// lengths[k] is sa_hsp's length; scores[k] is written only where the length
// is not -1 (search.c:113-116).
void gapless_seeds(const db_t* db, const seq_t* q, int n, const int* qoffsets,
                   const int64_t* positions, const options_t* opt, int* scores, int* lengths,
                   const char* who) {
    if (n <= 0) return;
    const ResidentText& rt = resident_text(db);
    for (int k = 0; k < n; ++k)
        if (db->SA && (positions[k] < 0 || (uint64_t)positions[k] >= db->Tsize)) {
            fprintf(stderr, "topaz_amd: %s: suffix-array slot %lld outside the text\n", who,
                    (long long)positions[k]);
            _exit(EXIT_FAILURE);
        }
    ThreadCtx& c = ctx();
    const int64_t zero64 = 0;
    const int32_t qlen = q->seq_len;
    const int8_t* mat = (const int8_t*)*opt->substitution;
    const bool same = c.g_text == rt.text && (size_t)qlen == c.g_query.size() &&
                      memcmp(c.g_query.data(), q->seq_buf.buf, (size_t)qlen) == 0 &&
                      memcmp(c.g_matrix, mat, 625) == 0;
    // query copies, score table and their offsets, uploaded once per query
    QueryCodes qc;
    size_t q_bytes = 0;
    if (!same) {
        if (gapless_query_codes(rt.text->codes, 1, q->seq_buf.buf, &zero64, &qlen, qc)) {
            fprintf(stderr, "topaz_amd: %s: %s\n", who, tpz_last_error());
            _exit(EXIT_FAILURE);
        }
        c.g_qr = (qc.fwd.size() + 15) & ~(size_t)15;
        c.g_tab = (c.g_qr + qc.rev.size() + 15) & ~(size_t)15;
        c.g_meta = c.g_tab + kTableRows * kTableCols * 2;
        q_bytes = c.g_meta + 32;
    }
    // pinned block: [query block, when it changed | seeds | results]; the
    // kernel reads the seeds from it and writes the results into it
    const size_t seed_bytes = (size_t)n * sizeof(SeedRec);
    char* h = c.pin.get(q_bytes + seed_bytes + (size_t)n * 8);
    if (!same) {
        const size_t gbytes = q_bytes;
        char* g = h;
        memset(g, 0, gbytes);
        memcpy(g, qc.fwd.data(), qc.fwd.size());
        memcpy(g + c.g_qr, qc.rev.data(), qc.rev.size());
        gapless_table(rt.text->codes, mat, (int16_t*)(g + c.g_tab));
        memcpy(g + c.g_meta, &zero64, 8);          // forward and reversed offsets
        memcpy(g + c.g_meta + 8, &zero64, 8);
        memcpy(g + c.g_meta + 16, &qlen, 4);
        char* gd = (char*)c.g_dev.get(gbytes);
        hip_or_die(hipMemcpyAsync(gd, g, gbytes, hipMemcpyHostToDevice, c.st), "upload");
        c.g_text = rt.text;
        c.g_query.assign(q->seq_buf.buf, q->seq_buf.buf + qlen);
        memcpy(c.g_matrix, mat, 625);
    }
    SeedRec* seeds = reinterpret_cast<SeedRec*>(h + q_bytes);
    for (int k = 0; k < n; ++k) seeds[k] = SeedRec{0, qoffsets[k], positions[k]};
    int32_t* res = reinterpret_cast<int32_t*>(h + q_bytes + seed_bytes);
    const char* gd = (const char*)c.g_dev.p;
    GaplessArgs a{};
    a.t_fwd = (const uint8_t*)rt.text->d_fwd.ptr;
    a.t_head = (const uint8_t*)rt.text->d_head.ptr;
    a.t_size = (int64_t)db->Tsize;
    a.sa = rt.text->has_sa ? (const int64_t*)rt.text->d_sa.ptr : nullptr;
    a.q_fwd = (const uint8_t*)gd;
    a.q_rev = (const uint8_t*)(gd + c.g_qr);
    a.qf_off = (const int64_t*)(gd + c.g_meta);
    a.qr_off = (const int64_t*)(gd + c.g_meta + 8);
    a.q_len = (const int32_t*)(gd + c.g_meta + 16);
    a.n_queries = 1;
    a.seeds = seeds;
    a.n_seeds = n;
    a.table = (const int16_t*)(gd + c.g_tab);
    a.out_score = res;
    a.out_length = res + n;
    const int grid = std::max(1, std::min((n + 255) / 256, std::max(num_cu(), 1) * 32));
    hip_or_die(tpz_launch_gapless(&a, grid, c.st), "gapless");
    c.finish("gapless");
    for (int k = 0; k < n; ++k) {
        lengths[k] = res[n + k];
        if (res[n + k] != -1) scores[k] = res[k];
    }
}
}  // namespace

int sa_hsp(db_t* db, seq_t* q, int qoffset, int64_t position, options_t* opt,
           int* gapless_score, int* gapless_length) {
    gapless_seeds(db, q, 1, &qoffset, &position, opt, gapless_score, gapless_length, "sa_hsp");
    return 0;
}

// ---------------------------------------------------------------------------
// batching extension
// ---------------------------------------------------------------------------
namespace {
// Subjects of the hits as matrix indices (db_seq then aa_table2 of upper_aa,
// db.c:233-247, ssw_wrapper.c:188-201), back to back; the batch re-packs them.
struct HitSubjects {
    std::vector<int8_t> res;
    std::vector<int64_t> off;
    std::vector<int32_t> len;
};

static void gather_subjects(const db_t* db, hit_t** hits, const std::vector<int>& which,
                            HitSubjects& out) {
    const Alphabet& ab = alphabet();
    out.res.clear(); out.off.clear(); out.len.clear();
    for (int k : which) {
        const hit_t* h = hits[k];
        if (h->db_index >= db->Psize) {                              // query.c:311-314
            fprintf(stderr, "Error: protein at index %llu not found!",
                    (unsigned long long)h->db_index);
            _exit(EXIT_FAILURE);
        }
        const position_t& p = db->index[h->db_index];
        out.off.push_back((int64_t)out.res.size());
        int32_t n = 0;
        for (int i = 0; i < p.protein_len; ++i) {
            const char c = db->T[p.protein_start + i];
            if (c == '\n') continue;
            out.res.push_back(ab.idx(ab.upper[(unsigned char)c]));
            ++n;
        }
        out.len.push_back(n);
    }
}
}  // namespace

namespace {
// One serving loop (tpz_server_*) per database, matrix and gap penalties,
// holding every protein of the database on the device.  TOPAZ's worker
// threads call align_batch_using_ssw_profile at the same time
// (engine.c:423-457); the loop merges their requests into shared launches,
// where one query's hits alone would leave most of the GPU idle for the
// length of its longest alignment.
struct DbServer {
    const db_t* db;
    const char* T;
    uint64_t psize;
    int8_t matrix[625];
    int go, ge;
    tpz_server* srv;
};
std::mutex g_srv_mu;
std::vector<DbServer*> g_servers;

// Stops the loops' threads before the HIP runtime is torn down at exit.
void stop_servers() {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    for (auto* s : g_servers) tpz_server_destroy(s->srv);
    g_servers.clear();
}

tpz_server* db_server(const db_t* db, const int8_t* mat, int go, int ge) {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    for (auto* s : g_servers)
        if (s->db == db && s->T == db->T && s->psize == db->Psize && s->go == go && s->ge == ge &&
            memcmp(s->matrix, mat, 625) == 0)
            return s->srv;
    // every protein as db_seq gives it (db.c:233-247), as matrix indices
    const Alphabet& ab = alphabet();
    std::vector<int8_t> res;
    std::vector<int64_t> off(db->Psize);
    std::vector<int32_t> len(db->Psize);
    for (uint64_t p = 0; p < db->Psize; ++p) {
        const position_t& pos = db->index[p];
        off[p] = (int64_t)res.size();
        for (int i = 0; i < pos.protein_len; ++i) {
            const char ch = db->T[pos.protein_start + i];
            if (ch != '\n') res.push_back(ab.idx(ab.upper[(unsigned char)ch]));
        }
        len[p] = (int32_t)(res.size() - (size_t)off[p]);
    }
    tpz_server_config cfg{};
    // launches in flight: TOPAZ_AMD_INFLIGHT (1..8), default 2
    cfg.max_inflight = 2;
    if (const char* e = getenv("TOPAZ_AMD_INFLIGHT")) {
        const int v = atoi(e);
        if (v < 1 || v > 8) {
            fprintf(stderr, "topaz_amd: TOPAZ_AMD_INFLIGHT=%s is not 1..8\n", e);
            _exit(EXIT_FAILURE);
        }
        cfg.max_inflight = v;
    }
    cfg.waves_per_group = 8;     // a query's hits are a task or two: 8-wavefront staircases
    auto* s = new DbServer{db, db->T, db->Psize, {0}, go, ge, nullptr};
    memcpy(s->matrix, mat, 625);
    if (tpz_server_create(&s->srv, (int32_t)db->Psize, res.data(), off.data(), len.data(), mat,
                          go, ge, &cfg))
        die("align_batch_using_ssw_profile: serving loop");
    if (g_servers.empty()) atexit(stop_servers);
    g_servers.push_back(s);
    return s->srv;
}

// The first seeds of a search through the batch ABI build the serving loop
// of its database, so that the proteins are on the device before the first
// alignments come: one worker thread pays for the build while the others
// go on with their seeds, instead of all of them waiting at their first
// alignment call.
void prefetch_server(const db_t* db, const options_t* opt) {
    static std::atomic<const db_t*> last{nullptr};
    static std::mutex mu;
    static std::vector<const db_t*> claimed;
    if (last.load(std::memory_order_acquire) == db) return;
    const int go = opt->gap_open, ge = opt->gap_extend;
    if (!gaps_valid(go, ge)) return;                       // the alignment call reports it
    if (!db->index || db->Psize == 0) return;
    {
        // one thread builds; the others do not wait for it here
        std::lock_guard<std::mutex> lk(mu);
        if (std::find(claimed.begin(), claimed.end(), db) != claimed.end()) return;
        claimed.push_back(db);
    }
    (void)db_server(db, (const int8_t*)*opt->substitution, go, ge);
    last.store(db, std::memory_order_release);
}
}  // namespace

int align_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                  options_t* opt) {
    check_gaps(opt->gap_open, opt->gap_extend);
    std::vector<int> todo;
    for (int k = 0; k < n; ++k) {
        hit_t* h = hits[k];
        if (h->gapless_score < 40) {                                 // ssw_wrapper.c:171-176
            h->bitscore = 0;
            h->evalue = 1000;
            h->is_aligned = 1;
        } else {
            todo.push_back(k);
        }
    }
    if (!todo.empty()) {
        const int m = (int)todo.size();
        std::vector<int32_t> pq(m, 0), ps(m), score(m), ref_end(m), read_end(m);
        for (int k = 0; k < m; ++k) {
            const hit_t* h = hits[todo[k]];
            if (h->db_index >= db->Psize) {                          // query.c:311-314
                fprintf(stderr, "Error: protein at index %llu not found!",
                        (unsigned long long)h->db_index);
                _exit(EXIT_FAILURE);
            }
            ps[k] = (int32_t)h->db_index;
        }
        tpz_server* srv = db_server(db, profile->mat, opt->gap_open, opt->gap_extend);
        const int64_t qo = 0;
        const int32_t ql = profile->readLen;
        uint64_t ticket = 0;
        double latency_ms = 0;
        if (tpz_server_submit(srv, 1, profile->read, &qo, &ql, m, pq.data(), ps.data(), &ticket) ||
            tpz_server_wait(srv, ticket, score.data(), ref_end.data(), read_end.data(), &latency_ms))
            die("align_batch_using_ssw_profile");
        double correction, corrected_q, corrected_db;
        karlin(nullptr, opt, profile->readLen, &correction, &corrected_q, &corrected_db);
        for (int k = 0; k < m; ++k) {
            hit_t* h = hits[todo[k]];
            if (score[k] >= 32767) {                                 // ssw_wrapper.c:222-226
                h->rawscore = 32767;
                continue;
            }
            // the forward pass alone leaves the begins at -1 (ssw.c:886-893)
            h->sstart = -1;
            h->send = ref_end[k];
            h->qstart = -1;
            h->qend = read_end[k];
            h->rawscore = score[k];
            h->length = 0;
            h->gapopen = 0;
            h->mismatch = 0;
            h->identity = 0.0;
            h->bitscore = (opt->lambda * h->rawscore - log(opt->k));  // ssw_wrapper.c:239-245
            h->evalue = corrected_q * corrected_db * exp(-h->bitscore);
            h->evalue = h->evalue > 1.0e-180 ? h->evalue : 0.0;
            h->bitscore /= M_LN2;
            h->is_aligned = 1;
            h->effective_length = profile->readLen - correction;
        }
    }
    int count = 0;
    for (int k = 0; k < n; ++k) count += hits[k]->evalue < opt->evalue;   // query.c:338
    return count;
}

void stats_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                   options_t* opt) {
    check_gaps(opt->gap_open, opt->gap_extend);
    std::vector<int> todo;
    for (int k = 0; k < n; ++k)
        if (hits[k]->rawscore != 32767) todo.push_back(k);           // ssw_wrapper.c:136-138
    if (todo.empty()) return;
    HitSubjects sub;
    gather_subjects(db, hits, todo, sub);
    const int m = (int)todo.size();
    std::vector<int32_t> hq(m, 0), hs(m), re(m), fe(m), sc(m);
    for (int k = 0; k < m; ++k) {
        const hit_t* h = hits[todo[k]];
        hs[k] = k;
        re[k] = h->qend;
        fe[k] = h->send;
        sc[k] = h->rawscore;
    }
    const int64_t qo = 0;
    const int32_t ql = profile->readLen;
    StatsOutput o;
    stats_run(StatsInput{1, profile->read, &qo, &ql, m, sub.res.data(), sub.off.data(),
                         sub.len.data(), m, hq.data(), hs.data(), re.data(), fe.data(), sc.data(),
                         profile->mat, opt->gap_open, opt->gap_extend},
              o, "stats_batch_using_ssw_profile");
    for (int k = 0; k < m; ++k) {
        hit_t* h = hits[todo[k]];
        h->qstart = o.read_begin[k];
        h->sstart = o.ref_begin[k];
        h->length = o.len[k];
        h->gapopen = o.gapo[k];
        h->mismatch = o.mis[k];
        h->identity = stats_identity(o.ident[k], o.len[k]);
    }
}

}  // extern "C"

namespace {
// Protein of every 64th text position, for db_index: the protein holding
// position p is the one holding p & ~63 or one of the next few.  Built once
// per database; used only when the index tiles the text the way TOPAZ
// writes it (each protein's letters and its '\n', back to back from 0).
struct ProteinMap {
    const db_t* db = nullptr;
    const void* index = nullptr;
    uint64_t t_size = 0, p_size = 0;
    uint64_t print = 0;             // a few entries, against a reused address
    bool tiled = false;
    std::vector<uint32_t> block;
};

uint64_t index_print(const db_t* db) {
    uint64_t h = 0;
    if (db->Psize == 0 || !db->index) return h;
    for (uint64_t k : {(uint64_t)0, db->Psize / 2, db->Psize - 1})
        h = h * 1000003 + db->index[k].protein_start * 31 + (uint64_t)db->index[k].protein_len;
    return h;
}
std::mutex g_pmap_mu;
std::vector<ProteinMap*> g_pmaps;
std::atomic<const ProteinMap*> g_pmap_last{nullptr};

bool same_db(const ProteinMap* m, const db_t* db) {
    return m->db == db && m->index == db->index && m->t_size == db->Tsize &&
           m->p_size == db->Psize && m->print == index_print(db);
}

const ProteinMap& protein_map(const db_t* db) {
    // every hit of a search asks: the last map answers without the lock
    const ProteinMap* last = g_pmap_last.load(std::memory_order_acquire);
    if (last && same_db(last, db)) return *last;
    std::lock_guard<std::mutex> lk(g_pmap_mu);
    for (auto* m : g_pmaps)
        if (same_db(m, db)) {
            g_pmap_last.store(m, std::memory_order_release);
            return *m;
        }
    auto* m = new ProteinMap();
    m->db = db;
    m->index = db->index;
    m->t_size = db->Tsize;
    m->p_size = db->Psize;
    m->print = index_print(db);
    uint64_t next = 0;
    bool tiled = db->Psize > 0 && db->Psize < UINT32_MAX;
    for (uint64_t k = 0; tiled && k < db->Psize; ++k) {
        tiled = db->index[k].protein_start == next && db->index[k].protein_len >= 0;
        next = db->index[k].protein_start + (uint64_t)db->index[k].protein_len + 1;
    }
    m->tiled = tiled && next == db->Tsize;
    if (m->tiled) {
        m->block.resize((db->Tsize + 63) / 64);
        uint64_t k = 0;
        for (uint64_t b = 0; b < m->block.size(); ++b) {
            const uint64_t pos = b * 64;
            while (pos > db->index[k].protein_start + (uint64_t)db->index[k].protein_len) ++k;
            m->block[b] = (uint32_t)k;
        }
    }
    g_pmaps.push_back(m);
    g_pmap_last.store(m, std::memory_order_release);
    return *m;
}
}  // namespace

extern "C" {

int db_index(db_t* db, uint64_t pos, int64_t* index) {
    if (pos >= db->Tsize) return -1;                                 // db.c:180-182
    const ProteinMap& m = protein_map(db);
    if (m.tiled) {
        // the protein whose letters or '\n' hold pos (db.c:184-203 finds the
        // same one: in a tiled index it is unique)
        uint64_t k = m.block[pos >> 6];
        while (pos > db->index[k].protein_start + (uint64_t)db->index[k].protein_len) ++k;
        *index = (int64_t)k;
        return 0;
    }
    // any other index: the reference's binary search, step for step
    uint64_t i, half, size;
    for (i = 0, size = db->Psize, half = size >> 1; 0 < size; size = half, half >>= 1) {
        const position_t* t = db->index + i + half;
        if (pos < t->protein_start) continue;
        if (pos > t->protein_start + (uint64_t)t->protein_len) {
            i += half + 1;
            half -= (size & 1) ^ 1;
            continue;
        }
        break;
    }
    *index = (int64_t)(i + half);
    return 0;
}

}  // extern "C"

// sa_search (search.c:237-361) for n offsets of one query on the GPU: each
// offset's index, match count and last position, and sa_search's return
// value in status[k].  The pattern at offset o is bytes o .. view_len - 1 of
// the query's buffer.
static void search_offsets(db_t* db, seq_t* q, int64_t view_len, int n, const int* offsets,
                           int64_t* idx, int64_t* num_matches, int64_t* last_position,
                           int32_t* status) {
    for (int k = 0; k < n; ++k)
        if (offsets[k] < 0) {
            // the reference reads before the query's first byte
            fprintf(stderr, "topaz_amd: sa_search: negative query offset %d\n", offsets[k]);
            _exit(EXIT_FAILURE);
        }
    const ResidentText& rt = resident_text(db);
    const char* raw = resident_raw(db);
    ThreadCtx& c = ctx();
    // pinned block: [query bytes | offsets | idx, num_matches, last_position | status]
    const size_t qb = ((size_t)view_len + 16) & ~(size_t)15;
    const size_t ob = ((size_t)n * 4 + 15) & ~(size_t)15;
    char* h = c.pin.get(qb + ob + (size_t)n * 24 + (size_t)n * 4);
    if (view_len > 0) memcpy(h, q->seq_buf.buf, (size_t)view_len);
    int32_t* ho = reinterpret_cast<int32_t*>(h + qb);
    for (int k = 0; k < n; ++k) ho[k] = offsets[k];
    int64_t* hr = reinterpret_cast<int64_t*>(h + qb + ob);
    int32_t* hs = reinterpret_cast<int32_t*>(h + qb + ob + (size_t)n * 24);
    SaSearchArgs a{};
    a.text = raw;
    a.t_size = (int64_t)db->Tsize;
    a.sa = (const int64_t*)rt.text->d_sa.ptr;
    a.query = h;
    a.q_len = view_len;
    a.offsets = ho;
    a.n = n;
    a.out = hr;
    a.status = hs;
    hip_or_die(tpz_launch_sa_search(&a, c.st), "sa_search");
    c.finish("sa_search");
    for (int k = 0; k < n; ++k) {
        idx[k] = hr[3 * k];
        status[k] = hs[k];
        if (hs[k] != 0) continue;
        if (num_matches) num_matches[k] = hr[3 * k + 1];
        if (last_position) last_position[k] = hr[3 * k + 2];
    }
}

extern "C" {

int sa_search_batch(db_t* db, seq_t* q, int n, const int* offsets, int64_t* idx,
                    int64_t* num_matches, int64_t* last_position) {
    if (n <= 0) return 0;
    if (!db->T || !db->SA || db->Tsize <= 0) {                     // search.c:274-276
        for (int k = 0; k < n; ++k) idx[k] = -1;
        return -1;
    }
    std::vector<int32_t> status((size_t)n);
    search_offsets(db, q, q->seq_len, n, offsets, idx, num_matches, last_position,
                   status.data());
    int rc = 0;
    for (int k = 0; k < n; ++k)
        if (status[k] != 0) rc = -1;
    return rc;
}

int sa_search(db_t* db, seq_t* q, int64_t offset, int64_t length, int64_t* idx,
              int64_t* num_matches, int64_t* last_position) {
    if (idx) *idx = -1;
    // Psize = length > 0 ? length : seq_len - offset (search.c:268-270)
    const int64_t psize = length > 0 ? length : (int64_t)q->seq_len - offset;
    if (!db->T || !db->SA || db->Tsize <= 0 || !q->seq_buf.buf || psize <= 0) return -1;
    if (offset < 0 || offset > INT32_MAX || offset + psize > (int64_t)q->seq_buf.len) {
        // the reference reads outside the query's buffer
        fprintf(stderr, "topaz_amd: sa_search: pattern [%lld, %lld) outside the query buffer "
                        "(%d bytes)\n", (long long)offset, (long long)(offset + psize),
                q->seq_buf.len);
        _exit(EXIT_FAILURE);
    }
    const int off = (int)offset;
    int64_t i = -1, nm = 0, lp = 0;
    int32_t st = 0;
    search_offsets(db, q, offset + psize, 1, &off, &i, &nm, &lp, &st);
    if (st != 0) return -1;
    if (idx) *idx = i;
    if (num_matches) *num_matches = nm;
    if (last_position) *last_position = lp;
    return 0;
}

void sa_hsp_batch(db_t* db, seq_t* q, int n, const int* qoffsets, const int64_t* positions,
                  options_t* opt, int* scores, int* lengths) {
    prefetch_server(db, opt);
    gapless_seeds(db, q, n, qoffsets, positions, opt, scores, lengths, "sa_hsp_batch");
}

}  // extern "C"
