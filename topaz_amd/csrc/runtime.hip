// Host runtime of libtopaz_amd: device selection, error reporting, and the
// batched Smith-Waterman driver behind the C ABI of include/topaz_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/topaz_amd.h"
#include "runtime.h"
#include "gapless_kernels.h"
#include "traceback_kernels.h"
#include "sw_kernels.h"
#include "sw_pair_kernels.h"
#include "sw_lane_kernels.h"

namespace tpz {

// ---------------------------------------------------------------------------
// errors and device
// ---------------------------------------------------------------------------
static thread_local std::string g_error;
static std::mutex g_init_mu;
static int g_device = -1;
static hipStream_t g_stream = nullptr;
static int g_num_cu = 0;

int fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return -1;
}

int check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail("%s: %s", what, hipGetErrorString(e));
}

int ensure_init() {
    if (g_device >= 0) return 0;
    return tpz_init(0);
}

hipStream_t stream() { return g_stream; }
int num_cu() { return g_num_cu; }
int device_id() { return g_device; }

DeviceBuffer::~DeviceBuffer() { reset(); }
void DeviceBuffer::reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
int DeviceBuffer::alloc(size_t n) {
    reset();
    if (n == 0) n = 4;
    if (check(hipMalloc(&ptr, n), "hipMalloc")) return -1;
    bytes = n;
    return 0;
}
int DeviceBuffer::upload(const void* src, size_t n) {
    if (alloc(n)) return -1;
    if (n && src)
        return check(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, g_stream), "upload");
    return 0;
}

// ---------------------------------------------------------------------------
// batched forward Smith-Waterman
// ---------------------------------------------------------------------------
static int64_t chunk_rows(int32_t m, int rows) {
    return (int64_t)(m + rows - 1) / rows * rows;
}

// ---------------------------------------------------------------------------
// host threads: one pool per process for every parallel loop of the library,
// so that concurrent calls (a search's workers, the server's scheduler) share
// the machine's cores instead of each starting threads of its own
// ---------------------------------------------------------------------------
int host_threads() {
    static const int n = [] {
        int v = (int)std::thread::hardware_concurrency();
        if (const char* e = getenv("TOPAZ_AMD_HOST_THREADS")) v = atoi(e);
        return std::max(1, std::min(v, 64));
    }();
    return n;
}

namespace {
class Pool {
public:
    explicit Pool(int n) {
        for (int k = 0; k < n; ++k) th_.emplace_back([this] { worker(); });
    }
    // fn(i) for i in [0, n), indices taken from a shared counter; the caller
    // works too (so a loop inside a loop cannot wait for itself).  Returns
    // when all are done.
    void run(int64_t n, const std::function<void(int64_t)>& fn) {
        if (n <= 0) return;
        auto job = std::make_shared<Job>();
        job->n = n;
        job->fn = &fn;
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(job);
        }
        cv_.notify_all();
        work(*job);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return job->done.load() == job->n; });
        jobs_.erase(std::remove(jobs_.begin(), jobs_.end(), job), jobs_.end());
    }

private:
    struct Job {
        int64_t n = 0;
        const std::function<void(int64_t)>* fn = nullptr;
        std::atomic<int64_t> next{0}, done{0};
    };
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::vector<std::shared_ptr<Job>> jobs_;

    void work(Job& j) {
        int64_t i;
        while ((i = j.next.fetch_add(1)) < j.n) {
            (*j.fn)(i);
            if (j.done.fetch_add(1) + 1 == j.n) {
                std::lock_guard<std::mutex> lk(mu_);
                done_cv_.notify_all();
            }
        }
    }
    // the threads live until the process ends
    void worker() {
        for (;;) {
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] {
                    for (auto& x : jobs_)
                        if (x->next.load() < x->n) return true;
                    return false;
                });
                for (auto& x : jobs_)
                    if (x->next.load() < x->n) { j = x; break; }
            }
            if (j) work(*j);
        }
    }
};

Pool& pool() {
    static Pool* p = new Pool(host_threads() - 1);
    return *p;
}
}  // namespace

void parallel_for_each(int64_t n, int threads, const std::function<void(int64_t)>& fn) {
    if (n <= 0) return;
    if (threads <= 1 || n == 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    pool().run(n, fn);
}

void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& range_fn) {
    if (n <= 0) return;
    const int t = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n));
    if (t == 1) {
        range_fn(0, n);
        return;
    }
    pool().run(t, [&](int64_t k) { range_fn(n * k / t, n * (k + 1) / t); });
}

int DeviceBuffer::upload_large(const void* src, size_t n) {
    if (alloc(n)) return -1;
    if (!n || !src) return 0;
    const uint8_t* s = static_cast<const uint8_t*>(src);
    return stage_to_device(ptr, n, [s](uint8_t* out, size_t b, size_t e) { memcpy(out, s + b, e - b); });
}

namespace {
constexpr int kStageBuffers = 3;
constexpr size_t kStageBytes = (size_t)128 << 20;
struct Staging {
    std::mutex mu;
    uint8_t* buf[kStageBuffers] = {};
    hipEvent_t ev[kStageBuffers] = {};
    bool ready = false;
};
Staging& staging() {
    static Staging* s = new Staging();   // lives until the process ends
    return *s;
}
}  // namespace

int stage_to_device(void* dst, size_t n,
                    const std::function<void(uint8_t* out, size_t begin, size_t end)>& fill) {
    if (n == 0) return 0;
    Staging& st = staging();
    std::lock_guard<std::mutex> lk(st.mu);
    if (!st.ready) {
        for (int k = 0; k < kStageBuffers; ++k) {
            if (check(hipHostMalloc((void**)&st.buf[k], kStageBytes, hipHostMallocDefault),
                      "hipHostMalloc (staging)") ||
                check(hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming), "hipEventCreate"))
                return -1;
        }
        st.ready = true;
    }
    const int threads = host_threads();
    uint8_t* d = static_cast<uint8_t*>(dst);
    size_t k = 0;
    for (size_t off = 0; off < n; off += kStageBytes, ++k) {
        const int b = (int)(k % kStageBuffers);
        // the copy that last read this buffer must have finished
        if (k >= (size_t)kStageBuffers && check(hipEventSynchronize(st.ev[b]), "staging wait"))
            return -1;
        const size_t len = std::min(kStageBytes, n - off);
        uint8_t* out = st.buf[b];
        // pieces of 1 MB over the host threads
        const int64_t pieces = (int64_t)((len + (1 << 20) - 1) >> 20);
        parallel_for(pieces, threads, [&](int64_t p0, int64_t p1) {
            const size_t a = (size_t)p0 << 20, e = std::min(len, (size_t)p1 << 20);
            if (a < e) fill(out + a, off + a, off + e);
        });
        if (check(hipMemcpyAsync(d + off, out, len, hipMemcpyHostToDevice, g_stream), "staged upload") ||
            check(hipEventRecord(st.ev[b], g_stream), "hipEventRecord"))
            return -1;
    }
    return check(hipStreamSynchronize(g_stream), "staged upload");
}


// A counting sort by query (pairs stay in pair order; per-thread counts over
// contiguous pieces of the pairs), then per query a stable LSD radix sort of
// its pairs by (longest subject - subject length): longest first, equal
// lengths in pair order.  Each subject length is read once.
std::vector<int2> pair_subjects(const int32_t* pair_subject, int64_t n_pairs, const int64_t* s_off,
                                const int32_t* s_len) {
    std::vector<int2> r((size_t)n_pairs);
    parallel_for(n_pairs, n_pairs >= (1 << 18) ? host_threads() : 1, [&](int64_t a, int64_t b) {
        for (int64_t p = a; p < b; ++p) {
            const int32_t k = pair_subject[p];
            r[(size_t)p] = make_int2((int32_t)(uint32_t)(s_off[k] >> 4), s_len[k]);
        }
    });
    return r;
}

QueryPairs pairs_by_query(const std::vector<int32_t>& pair_query, const int2* pair_sub,
                          int32_t n_queries) {
    QueryPairs r;
    const int64_t n = (int64_t)pair_query.size();
    const int threads = n >= (1 << 18) ? host_threads() : 1;
    // pieces of at least 64 K pairs, and per-piece counts of at most ~4 words per pair
    const int pieces = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)threads, n >> 16, 4 * n / std::max<int64_t>(1, n_queries)}));
    r.begin.assign((size_t)n_queries + 1, 0);
    r.pair.resize((size_t)n);
    r.len.resize((size_t)n);
    // cnt[j][q]: pairs of query q in piece j, then where piece j's pairs of q go
    std::vector<std::vector<int64_t>> cnt((size_t)pieces, std::vector<int64_t>((size_t)n_queries, 0));
    parallel_for(pieces, pieces, [&](int64_t j0, int64_t j1) {
        for (int64_t j = j0; j < j1; ++j)
            for (int64_t p = n * j / pieces, e = n * (j + 1) / pieces; p < e; ++p)
                ++cnt[(size_t)j][(size_t)pair_query[p]];
    });
    int64_t at = 0;
    for (int32_t q = 0; q < n_queries; ++q) {
        r.begin[q] = at;
        for (int j = 0; j < pieces; ++j) {
            const int64_t c = cnt[(size_t)j][(size_t)q];
            cnt[(size_t)j][(size_t)q] = at;
            at += c;
        }
    }
    r.begin[n_queries] = at;
    parallel_for(pieces, pieces, [&](int64_t j0, int64_t j1) {
        for (int64_t j = j0; j < j1; ++j) {
            std::vector<int64_t>& w = cnt[(size_t)j];
            for (int64_t p = n * j / pieces, e = n * (j + 1) / pieces; p < e; ++p)
                r.pair[w[(size_t)pair_query[p]]++] = (int32_t)p;
        }
    });
    parallel_for_each(n_queries, threads, [&](int64_t q) {
        const int64_t b = r.begin[q], m = r.begin[q + 1] - b;
        if (m <= 1) {
            if (m == 1) r.len[b] = pair_sub[r.pair[b]].y;
            return;
        }
        int32_t* pr = r.pair.data() + b;
        int32_t* ln = r.len.data() + b;
        int32_t longest = 0;
        for (int64_t i = 0; i < m; ++i) longest = std::max(longest, ln[i] = pair_sub[pr[i]].y);
        // digits of (longest - length), least significant first, at most 12 bits each
        int bits = 0;
        while (bits < 31 && ((int64_t)longest >> bits) != 0) ++bits;
        if (bits == 0) return;                 // every subject empty: pair order
        const int passes = (bits + 11) / 12, width = (bits + passes - 1) / passes;
        thread_local std::vector<int32_t> tp, tl;
        thread_local std::vector<int64_t> count;
        tp.resize((size_t)m);
        tl.resize((size_t)m);
        count.resize((size_t)1 << width);
        int32_t *sp = pr, *sl = ln, *dp = tp.data(), *dl = tl.data();
        for (int pass = 0; pass < passes; ++pass) {
            const int shift = pass * width;
            const int32_t mask = (1 << width) - 1;
            std::fill(count.begin(), count.end(), 0);
            for (int64_t i = 0; i < m; ++i) ++count[(size_t)(((longest - sl[i]) >> shift) & mask)];
            int64_t sum = 0;
            for (auto& c : count) {
                const int64_t v = c;
                c = sum;
                sum += v;
            }
            for (int64_t i = 0; i < m; ++i) {
                const int64_t o = count[(size_t)(((longest - sl[i]) >> shift) & mask)]++;
                dp[o] = sp[i];
                dl[o] = sl[i];
            }
            std::swap(sp, dp);
            std::swap(sl, dl);
        }
        if (sp != pr) {
            std::copy(sp, sp + m, pr);
            std::copy(sl, sl + m, ln);
        }
    });
    return r;
}

// Positions of a QueryPairs (each query's pairs, longest first) not taken
// yet: untaken(i) is the first untaken position >= i, untaken_back(i) the
// last one <= i (-1: none).  Path halving; pv is shifted by one.  Queries own
// disjoint ranges, so groups of queries can be planned at once.
struct Untaken {
    std::unique_ptr<int32_t[]> nx, pv;   // positions fit int32 (pair counts are int32)
    Untaken(int64_t n, int threads) : nx(new int32_t[(size_t)n + 1]), pv(new int32_t[(size_t)n + 1]) {
        parallel_for(n + 1, threads, [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) nx[i] = pv[i] = (int32_t)i;
        });
    }
    int64_t next(int64_t i) {
        int32_t* x = nx.get();
        while (x[i] != i) { x[i] = x[x[i]]; i = x[i]; }
        return i;
    }
    int64_t back(int64_t i) {
        int32_t* x = pv.get();
        int64_t k = i + 1;
        while (x[k] != k) { x[k] = x[x[k]]; k = x[k]; }
        return k - 1;
    }
    void take(int64_t i) { nx[i] = (int32_t)(i + 1); pv[i + 1] = (int32_t)i; }
};

// The tasks of one group of queries with the same chunk count (see
// build_packed_tasks), appended to t.
static void plan_group(const QueryPairs& per, const int32_t* qs, size_t nqs, int64_t task_rows,
                       int nq_max, bool fill, Untaken& U, Tasks& t) {
    std::vector<int64_t> head(nqs);
    std::vector<int32_t> id(qs, qs + nqs);
    int64_t n_pairs = 0;
    for (size_t k = 0; k < nqs; ++k) {
        head[k] = per.begin[id[k]];
        n_pairs += per.count(id[k]);
    }
    const size_t max_tasks = (size_t)(n_pairs / kTaskPairs) + nqs;
    t.index.reserve(t.index.size() + (size_t)n_pairs);
    t.next.reserve(t.next.size() + (size_t)n_pairs);
    t.begin.reserve(t.begin.size() + max_tasks);
    t.query.reserve(t.query.size() + max_tasks * kTaskQueries);
    t.cols.reserve(t.cols.size() + max_tasks);
    t.cost.reserve(t.cost.size() + max_tasks);
    // local index k: is there a remaining pair, and the longest's length
    auto has = [&](size_t k) { return (head[k] = U.next(head[k])) < per.begin[id[k] + 1]; };
    auto shortest = [&](size_t k) {
        const int64_t i = U.back(per.begin[id[k] + 1] - 1);
        return i >= per.begin[id[k]] ? per.len[i] : INT32_MAX;
    };
    // max-heap on (longest remaining subject, then lower query number)
    auto below = [&](size_t x, size_t y) {
        const int32_t lx = per.len[head[x]], ly = per.len[head[y]];
        return lx != ly ? lx < ly : id[x] > id[y];
    };
    std::vector<size_t> heap(nqs);
    std::iota(heap.begin(), heap.end(), (size_t)0);
    std::make_heap(heap.begin(), heap.end(), below);
    std::vector<int32_t> load, last;
    load.reserve(kTaskPairs);
    last.reserve(kTaskPairs);
    while (!heap.empty()) {
        size_t sel[kTaskQueries];
        int32_t selq[kTaskQueries];
        int ns = 0;
        while (ns < nq_max && !heap.empty()) {
            std::pop_heap(heap.begin(), heap.end(), below);
            sel[ns] = heap.back();
            selq[ns] = id[heap.back()];
            ++ns;
            heap.pop_back();
        }
        t.open(selq, ns);
        load.clear();
        last.clear();
        // the seeds: a merge of the selected queries' longest remaining
        // pairs; only the query just taken from moves on
        int32_t hl[kTaskQueries];
        for (int k = 0; k < ns; ++k) hl[k] = has(sel[k]) ? per.len[head[sel[k]]] : -1;
        for (int taken = 0; taken < kTaskPairs; ++taken) {
            int best = -1;
            for (int k = 0; k < ns; ++k)
                if (hl[k] >= 0 && (best < 0 || hl[k] > hl[best] ||
                                   (hl[k] == hl[best] && selq[k] < selq[best])))
                    best = k;
            if (best < 0) break;
            const size_t b = sel[best];
            load.push_back(item_columns(per.len[head[b]]));
            last.push_back((int32_t)t.index.size());
            t.index.push_back(per.pair[head[b]]);
            t.next.push_back(-1);
            U.take(head[b]);
            hl[best] = has(b) ? per.len[head[b]] : -1;
        }
        const int32_t width = load.empty() ? 0 : load[0];
        if (fill) {
            // runs from the shortest (the last seeded) up take the longest
            // remaining pair that fits (a pair of length <= gap takes
            // item_columns <= gap columns: gaps are multiples of 4); a run
            // whose gap is below every query's shortest remaining pair takes
            // none, and neither do the runs after it (smaller gaps)
            int32_t least = INT32_MAX;
            for (int k = 0; k < ns; ++k) least = std::min(least, shortest(sel[k]));
            for (int lane = (int)load.size() - 1; lane >= 0; --lane) {
                if (width - load[lane] < 8 || width - load[lane] < least) break;
                for (;;) {
                    const int32_t gap = width - load[lane];
                    if (gap < 8 || gap < least) break;
                    int64_t pick = -1;
                    for (int k = 0; k < ns; ++k) {
                        const int64_t b = per.begin[selq[k]], e = per.begin[selq[k] + 1];
                        // first position of the query whose length is <= gap
                        const int64_t f = std::lower_bound(per.len.begin() + b, per.len.begin() + e, gap,
                                                           std::greater<int32_t>()) - per.len.begin();
                        const int64_t i = f < e ? U.next(f) : e;
                        if (i < e && (pick < 0 || per.len[i] > per.len[pick])) pick = i;
                    }
                    if (pick < 0) break;
                    const int32_t e = (int32_t)t.index.size();
                    t.index.push_back(per.pair[pick]);
                    t.next.push_back(-1);
                    t.next[last[lane]] = e;
                    last[lane] = e;
                    load[lane] += item_columns(per.len[pick]);
                    U.take(pick);
                    least = INT32_MAX;
                    for (int k = 0; k < ns; ++k) least = std::min(least, shortest(sel[k]));
                }
            }
        }
        t.cols.back() = width;
        t.cost.back() = task_rows * width;
        for (int k = 0; k < ns; ++k) {
            if (has(sel[k])) {
                heap.push_back(sel[k]);
                std::push_heap(heap.begin(), heap.end(), below);
            }
        }
    }
}

// Tasks of up to nq (<= 4) queries (the persistent schedule).  Queries with
// the same number of row chunks are packed together: repeatedly the nq
// queries whose longest remaining subjects are longest give their longest
// remaining pairs, longest first, until the task holds 128: one per lane
// half, the first of its run.  A query's long subjects then share a
// wavefront with other queries' long subjects instead of its own short ones.
// The task's width is its longest pair; with fill, every run, from the
// shortest up, then takes the longest remaining pair of the same queries
// that fits in what is left of that width, while one does, so the lane
// halves walk pairs instead of padding columns.  Groups of queries are
// planned on host threads at once (groups of more than 64 queries in parts
// of 64) and their tasks concatenated.
void build_packed_tasks(const QueryPairs& per, const std::vector<int32_t>& q_len, int rows,
                        int nq_max, Tasks& t, bool fill) {
    t.clear();
    const int32_t nq = (int32_t)per.begin.size() - 1;
    std::vector<int32_t> order;
    for (int32_t q = 0; q < nq; ++q)
        if (per.count(q) > 0) order.push_back(q);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return chunk_rows(q_len[x], rows) < chunk_rows(q_len[y], rows);
    });
    // parts: runs of order with one chunk count, at most 64 queries each
    constexpr size_t kPart = 64;
    std::vector<size_t> part;
    for (size_t g0 = 0; g0 < order.size();) {
        const int64_t r = chunk_rows(q_len[order[g0]], rows);
        size_t g1 = g0;
        while (g1 < order.size() && g1 - g0 < kPart && chunk_rows(q_len[order[g1]], rows) == r) ++g1;
        part.push_back(g0);
        g0 = g1;
    }
    part.push_back(order.size());
    const int64_t n_parts = (int64_t)part.size() - 1;
    const int threads = per.begin.back() >= (1 << 18) ? host_threads() : 1;
    Untaken U(per.begin.back(), threads);
    std::vector<Tasks> sub((size_t)n_parts);
    // parts with the most pairs first, each to the next free thread
    std::vector<int64_t> part_pairs((size_t)n_parts, 0), by_size((size_t)n_parts);
    for (int64_t k = 0; k < n_parts; ++k)
        for (size_t g = part[k]; g < part[k + 1]; ++g) part_pairs[k] += per.count(order[g]);
    std::iota(by_size.begin(), by_size.end(), (int64_t)0);
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int64_t x, int64_t y) { return part_pairs[x] > part_pairs[y]; });
    parallel_for_each(n_parts, threads, [&](int64_t i) {
        const int64_t k = by_size[i];
        plan_group(per, order.data() + part[k], part[k + 1] - part[k],
                   chunk_rows(q_len[order[part[k]]], rows), nq_max, fill, U, sub[k]);
    });
    // concatenate: each part's tasks and entries at its offsets
    std::vector<int64_t> e_at((size_t)n_parts + 1, 0), t_at((size_t)n_parts + 1, 0);
    for (int64_t k = 0; k < n_parts; ++k) {
        e_at[k + 1] = e_at[k] + (int64_t)sub[k].index.size();
        t_at[k + 1] = t_at[k] + sub[k].count();
    }
    const size_t n_tasks = (size_t)t_at[n_parts], n_entries = (size_t)e_at[n_parts];
    t.begin.resize(n_tasks + 1);
    t.cols.resize(n_tasks);
    t.cost.resize(n_tasks);
    t.query.resize(n_tasks * kTaskQueries);
    t.index.resize(n_entries);
    t.next.resize(n_entries);
    parallel_for(n_parts, threads, [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k) {
            const Tasks& g = sub[k];
            const int32_t eb = (int32_t)e_at[k];
            const size_t tb = (size_t)t_at[k];
            for (int32_t j = 0; j < g.count(); ++j) {
                t.begin[tb + j] = eb + g.begin[j];
                t.cols[tb + j] = g.cols[j];
                t.cost[tb + j] = g.cost[j];
            }
            std::copy(g.query.begin(), g.query.end(), t.query.begin() + tb * kTaskQueries);
            std::copy(g.index.begin(), g.index.end(), t.index.begin() + eb);
            for (size_t x = 0; x < g.next.size(); ++x)
                t.next[eb + x] = g.next[x] >= 0 ? g.next[x] + eb : -1;
        }
    });
    t.begin[n_tasks] = (int32_t)n_entries;
}

void build_entry_records(const Tasks& t, const int32_t* pair_query, const int2* pair_sub,
                         std::vector<int4>& rec) {
    rec.resize(t.index.size());
    const int32_t n_tasks = t.count();
    const int threads = t.index.size() >= ((size_t)1 << 20) ? host_threads() : 1;
    parallel_for(n_tasks, threads, [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k) {
            const int32_t* tq = t.query.data() + (size_t)kTaskQueries * k;
            for (int32_t e = t.begin[k]; e < t.begin[k + 1]; ++e) {
                const int32_t p = t.index[e];
                const int32_t q = pair_query[p];
                int slot = 0;
                for (int j = 1; j < kTaskQueries; ++j)
                    if (tq[j] == q) slot = j;
                int4 r;
                r.x = pair_sub[p].x;
                r.y = pair_sub[p].y;
                r.z = t.next[e];
                r.w = slot;
                rec[e] = r;
            }
        }
    });
}

}  // namespace tpz

using namespace tpz;

// Everything that depends on the batch's configuration: tasks, work items,
// workspace and the captured launch graphs.
struct SwPlan {
    tpz_sw_config cfg{};
    DeviceTasks packed_tasks;                       // up to queries_per_task per task
    // work items (W task slots each) of the persistent schedule, by
    // decreasing cost
    std::vector<int32_t> items_persistent;
    DeviceBuffer d_items_persistent;
    // plain schedule: the pairs of each query (a launch per query, a
    // wavefront per pair), their workspace offsets when the query needs
    // more than one pass, and the workspace of the largest such launch
    std::vector<int32_t> plain_begin;               // per query, into plain_pairs
    DeviceBuffer d_plain_pairs, d_plain_ws_off, d_plain_ws;
    int32_t big_tasks = 0;
    size_t ws_slots = 0;
    int64_t ws_words_per_slot = 0;
    DeviceBuffer d_workspace;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    SwPlan() = default;
    SwPlan(const SwPlan&) = delete;
    SwPlan& operator=(const SwPlan&) = delete;
    ~SwPlan() {
        for (int k = 0; k < 2; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
        }
    }
};

struct tpz_sw_batch {
    int32_t n_queries = 0, n_subjects = 0, n_pairs = 0;
    int32_t gap_open = 11, gap_extend = 1, bias = 0;
    int32_t max_subject = 0;
    std::vector<int32_t> q_len, s_len, pair_query, pair_subject;
    std::vector<int64_t> s_off;                     // packed subject offsets (16-byte aligned)
    int64_t s_pad = 0;                              // offset of the padding after the last subject
    QueryPairs per_query;                           // pairs of each query, longest subject first
    std::vector<int2> pair_sub;                     // per pair: subject offset / 16, length
    DeviceBuffer d_q_res, d_q_off, d_q_len, d_s_res, d_s_off, d_s_len;
    DeviceBuffer d_pair_query, d_pair_subject, d_matrix;
    DeviceBuffer d_score, d_ref, d_read;            // results
    DeviceBuffer d_key;                             // per entry: best key of the forward kernel
    DeviceBuffer d_counter, d_error;
    std::unique_ptr<SwPlan> plan;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms = 0;
    // gap_open <= gap_extend: the reference's lane recurrence (sw_lane_kernel)
    // in place of the textbook one, with its own workspace
    bool lanes = false;
    DeviceBuffer d_lane_ws;
    int lane_blocks = 0;
    int64_t lane_ws_rows = 0;
};

namespace tpz {
// Work items of the persistent schedule.  A task whose single-wavefront
// time would exceed big_fraction of one wavefront's fair share of the batch
// is worked by a whole workgroup (chunks interleaved over 4 wavefronts);
// the others go four to an item, one per wavefront, neighbours in cost
// order.  Items run most expensive first.
void build_items(const Tasks& t, int n_slots, double big_fraction, int waves,
                        std::vector<int32_t>& items, int32_t& n_big) {
    const size_t n = t.cost.size();
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return t.cost[x] > t.cost[y]; });
    double total = 0;
    for (int64_t c : t.cost) total += (double)c;
    const double share = total / std::max(1, n_slots * waves);
    std::vector<std::pair<double, std::vector<int32_t>>> list;
    n_big = 0;
    size_t k = 0;
    while (k < n && t.cost[order[k]] > big_fraction * share && t.cost[order[k]] > 0) {
        std::vector<int32_t> it(waves, -1);
        it[0] = order[k];
        it[1] = -2;
        list.push_back({t.cost[order[k]] / (double)waves, it});
        ++n_big;
        ++k;
    }
    while (k < n) {
        std::vector<int32_t> it(waves, -1);
        const double c0 = (double)t.cost[order[k]];
        for (int w = 0; w < waves && k < n; ++w) it[w] = order[k++];
        list.push_back({c0, it});
    }
    std::stable_sort(list.begin(), list.end(),
                     [](const auto& x, const auto& y) { return x.first > y.first; });
    items.clear();
    for (auto& e : list) items.insert(items.end(), e.second.begin(), e.second.end());
}

static tpz_sw_config default_config() {
    tpz_sw_config c{};
    c.rows_per_chunk = 32;
    c.groups_per_cu = kGroupsPerCu;
    c.queries_per_task = kTaskQueries;
    c.use_graph = 0;
    c.big_fraction = 0.5;
    c.waves_per_group = kGroupWaves;
    return c;
}

static int check_config(const tpz_sw_config& c) {
    if (c.rows_per_chunk != 16 && c.rows_per_chunk != 32)
        return fail("rows_per_chunk must be 16 or 32 (got %d)", c.rows_per_chunk);
    if (c.waves_per_group != 4 && c.waves_per_group != 8)
        return fail("waves_per_group must be 4 or 8 (got %d)", c.waves_per_group);
    // resident workgroups per CU allowed by the kernel's registers and LDS
    const int max_groups =
        std::max(1, tpz_sw_forward_groups_per_cu(c.rows_per_chunk, c.waves_per_group));
    if (c.groups_per_cu < 1 || c.groups_per_cu > max_groups)
        return fail("groups_per_cu must be 1..%d for %d rows and %d waves (got %d)", max_groups,
                    c.rows_per_chunk, c.waves_per_group, c.groups_per_cu);
    if (c.queries_per_task != 1 && c.queries_per_task != 2 && c.queries_per_task != 4)
        return fail("queries_per_task must be 1, 2 or 4 (got %d)", c.queries_per_task);
    if (!(c.big_fraction >= 0)) return fail("big_fraction must be >= 0");
    return 0;
}

// Tasks, work items and workspace of the batch for configuration c.
static int build_plan(tpz_sw_batch* b, const tpz_sw_config& c) {
    if (check_config(c)) return -1;
    auto plan = std::make_unique<SwPlan>();
    plan->cfg = c;
    build_packed_tasks(b->per_query, b->q_len, c.rows_per_chunk,
                       c.queries_per_task, plan->packed_tasks.host);
    std::vector<int4> rec;
    build_entry_records(plan->packed_tasks.host, b->pair_query.data(), b->pair_sub.data(), rec);
    int rc = plan->packed_tasks.upload(rec);
    const int W = c.waves_per_group;
    {
        // plain: each query's pairs in one list; a query longer than one
        // pass of the pair kernel needs 2 words per subject column per pair
        std::vector<int32_t> pairs;
        std::vector<int64_t> ws_off;
        int64_t ws_max = 0;
        for (int32_t q = 0; q < b->n_queries; ++q) {
            plan->plain_begin.push_back((int32_t)pairs.size());
            const bool passes = b->q_len[q] > 64 * tpz_sw_pair_rows(b->q_len[q]);
            int64_t ws = 0;
            for (int64_t i = b->per_query.begin[q]; i < b->per_query.begin[q + 1]; ++i) {
                const int32_t p = b->per_query.pair[i];
                pairs.push_back(p);
                ws_off.push_back(ws);
                if (passes) ws += 2 * (int64_t)b->s_len[b->pair_subject[p]];
            }
            ws_max = std::max(ws_max, ws);
        }
        plan->plain_begin.push_back((int32_t)pairs.size());
        rc |= plan->d_plain_pairs.upload(pairs.data(), pairs.size() * 4);
        rc |= plan->d_plain_ws_off.upload(ws_off.data(), ws_off.size() * 8);
        rc |= plan->d_plain_ws.alloc((size_t)ws_max * 4 + 16);
    }
    const size_t persistent_slots = (size_t)std::max(num_cu(), 1) * c.groups_per_cu;
    build_items(plan->packed_tasks.host, (int)persistent_slots, c.big_fraction, W,
                plan->items_persistent, plan->big_tasks);
    rc |= plan->d_items_persistent.upload(plan->items_persistent.data(),
                                          plan->items_persistent.size() * 4);
    // workspace: per workgroup, one boundary buffer (bottom H and F of a
    // chunk: 2 words per column and lane) per wavefront
    plan->ws_words_per_slot = (int64_t)W * 2 * (item_columns(b->max_subject) + 4) * 64;
    const size_t n_items_p = plan->items_persistent.size() / W;
    plan->ws_slots = std::max<size_t>(std::min(persistent_slots, n_items_p), 1);
    rc |= plan->d_workspace.alloc(plan->ws_slots * (size_t)plan->ws_words_per_slot * 4);
    if (rc) return -1;
    // the old plan's buffers may still be read by queued work
    if (check(hipStreamSynchronize(stream()), "plan")) return -1;
    b->plan = std::move(plan);
    return 0;
}
}  // namespace tpz

extern "C" {

int tpz_init(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device >= 0) return g_device == device ? 0 : fail("already bound to device %d", g_device);
    int n = 0;
    if (check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return -1;
    if (n <= 0) return fail("no HIP device available");
    if (device < 0 || device >= n) return fail("device %d out of range (%d devices)", device, n);
    if (check(hipSetDevice(device), "hipSetDevice")) return -1;
    hipDeviceProp_t prop;
    if (check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties")) return -1;
    g_num_cu = prop.multiProcessorCount;
    if (check(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate")) return -1;
    g_device = device;
    return 0;
}

const char* tpz_last_error(void) { return g_error.c_str(); }
const char* tpz_version(void) { return "topaz_amd 0.1 (gfx950)"; }

int tpz_sw_batch_create(tpz_sw_batch** out, int32_t n_queries, const int8_t* q_res,
                        const int64_t* q_off, const int32_t* q_len, int32_t n_subjects,
                        const int8_t* s_res, const int64_t* s_off, const int32_t* s_len,
                        int32_t n_pairs, const int32_t* pair_query,
                        const int32_t* pair_subject, const int8_t* matrix,
                        int32_t gap_open, int32_t gap_extend) {
    if (!out) return fail("tpz_sw_batch_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    if (n_queries < 0 || n_subjects < 0 || n_pairs < 0) return fail("negative counts");
    if (!gaps_valid(gap_open, gap_extend))
        return fail("gap penalties must lie in [0, 255] (got %d/%d)", gap_open, gap_extend);
    if (!matrix) return fail("null matrix");
    int32_t minw = 0;
    for (int i = 0; i < 625; ++i) {
        minw = std::min<int32_t>(minw, matrix[i]);
    }
    auto b = new tpz_sw_batch();
    b->n_queries = n_queries; b->n_subjects = n_subjects; b->n_pairs = n_pairs;
    b->gap_open = gap_open; b->gap_extend = gap_extend; b->bias = -minw;
    b->q_len.assign(q_len, q_len + n_queries);
    b->s_len.assign(s_len, s_len + n_subjects);
    b->pair_query.assign(pair_query, pair_query + n_pairs);
    b->pair_subject.assign(pair_subject, pair_subject + n_pairs);
    for (int32_t k = 0; k < n_queries; ++k) {
        if (q_len[k] < 0) { delete b; return fail("query %d has negative length", k); }
        if (q_len[k] > kMaxSwLength) {
            delete b;
            return fail("query %d is longer than %d residues", k, kMaxSwLength);
        }
    }
    for (int32_t p = 0; p < n_pairs; ++p) {
        if (pair_query[p] < 0 || pair_query[p] >= n_queries ||
            pair_subject[p] < 0 || pair_subject[p] >= n_subjects) {
            delete b;
            return fail("pair %d references query %d / subject %d out of range", p,
                        pair_query[p], pair_subject[p]);
        }
    }
    // queries: contiguous copy with their own offsets
    std::vector<int64_t> qo(n_queries);
    int64_t qtot = 0;
    for (int32_t k = 0; k < n_queries; ++k) { qo[k] = qtot; qtot += q_len[k]; }
    std::vector<int8_t> qr((size_t)qtot + 64);
    for (int32_t k = 0; k < n_queries; ++k) {
        for (int32_t j = 0; j < q_len[k]; ++j) {
            int8_t r = q_res[q_off[k] + j];
            if (r < 0 || r > 24) { delete b; return fail("query %d residue %d out of range", k, r); }
            qr[qo[k] + j] = r;
        }
    }
    // subjects: each starts on a 16-byte boundary, padded with residue 25
    std::vector<int64_t> so(n_subjects);
    int64_t stot = 0;
    for (int32_t k = 0; k < n_subjects; ++k) {
        if (s_len[k] < 0) { delete b; return fail("subject %d has negative length", k); }
        if (s_len[k] > kMaxSwLength) {
            delete b;
            return fail("subject %d is longer than %d residues", k, kMaxSwLength);
        }
        so[k] = stot;
        stot += (s_len[k] + 15) & ~15;
        b->max_subject = std::max(b->max_subject, s_len[k]);
    }
    b->s_pad = stot;
    std::vector<int8_t> sr((size_t)stot + 64, 25);
    for (int32_t k = 0; k < n_subjects; ++k) {
        for (int32_t j = 0; j < s_len[k]; ++j) {
            int8_t r = s_res[s_off[k] + j];
            if (r < 0 || r > 24) { delete b; return fail("subject %d residue %d out of range", k, r); }
            sr[so[k] + j] = r;
        }
    }
    int rc = 0;
    rc |= b->d_q_res.upload(qr.data(), qr.size());
    rc |= b->d_q_off.upload(qo.data(), qo.size() * 8);
    rc |= b->d_q_len.upload(q_len, (size_t)n_queries * 4);
    rc |= b->d_s_res.upload(sr.data(), sr.size());
    rc |= b->d_s_off.upload(so.data(), so.size() * 8);
    rc |= b->d_s_len.upload(s_len, (size_t)n_subjects * 4);
    rc |= b->d_pair_query.upload(pair_query, (size_t)n_pairs * 4);
    rc |= b->d_pair_subject.upload(pair_subject, (size_t)n_pairs * 4);
    rc |= b->d_matrix.upload(matrix, 625);
    rc |= b->d_key.alloc((size_t)n_pairs * 8 + 8);
    b->s_off = std::move(so);
    rc |= b->d_score.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_ref.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_read.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_counter.alloc(64);
    if (rc) { delete b; return -1; }

    rc |= b->d_error.alloc(64);
    if (!rc) rc |= check(hipMemsetAsync(b->d_error.ptr, 0, 64, stream()), "memset");
    b->pair_sub = pair_subjects(b->pair_subject.data(), n_pairs, b->s_off.data(), b->s_len.data());
    b->per_query = pairs_by_query(b->pair_query, b->pair_sub.data(), n_queries);
    if (!textbook_gaps(gap_open, gap_extend) && n_pairs > 0) {
        b->lanes = true;
        int32_t max_q = 0;
        for (int32_t k = 0; k < n_queries; ++k) max_q = std::max(max_q, q_len[k]);
        b->lane_ws_rows = lane_rows(max_q);
        b->lane_blocks = (int)std::min<int64_t>((n_pairs + kLaneGroupsPerBlock - 1) / kLaneGroupsPerBlock,
                                                (int64_t)std::max(num_cu(), 1) * 16);
        rc |= b->d_lane_ws.alloc((size_t)b->lane_blocks * kLaneGroupsPerBlock * 3 *
                                 (size_t)b->lane_ws_rows * 4 + 64);
    }
    for (auto& e : b->ev) rc |= check(hipEventCreate(&e), "hipEventCreate");
    tpz_sw_config cfg = default_config();
    if (!rc) rc |= build_plan(b, cfg);
    rc |= check(hipStreamSynchronize(stream()), "upload sync");
    if (rc) { tpz_sw_batch_destroy(b); return -1; }
    *out = b;
    return 0;
}

static SwKernelArgs make_args(tpz_sw_batch* b) {
    SwKernelArgs a{};
    a.q_res = (const int8_t*)b->d_q_res.ptr;
    a.q_off = (const int64_t*)b->d_q_off.ptr;
    a.q_len = (const int32_t*)b->d_q_len.ptr;
    a.s_res = (const int8_t*)b->d_s_res.ptr;
    a.s_pad = b->s_pad;
    a.s_off = (const int64_t*)b->d_s_off.ptr;
    a.s_len = (const int32_t*)b->d_s_len.ptr;
    a.pair_query = (const int32_t*)b->d_pair_query.ptr;
    a.pair_subject = (const int32_t*)b->d_pair_subject.ptr;
    a.out_key = (unsigned long long*)b->d_key.ptr;
    a.matrix = (const int8_t*)b->d_matrix.ptr;
    a.gap_open = b->gap_open;
    a.gap_extend = b->gap_extend;
    a.workspace = (uint32_t*)b->plan->d_workspace.ptr;
    a.ws_words_per_slot = b->plan->ws_words_per_slot;
    a.out_score = (int32_t*)b->d_score.ptr;
    a.out_ref = (int32_t*)b->d_ref.ptr;
    a.out_read = (int32_t*)b->d_read.ptr;
    a.error = (int32_t*)b->d_error.ptr;
    return a;
}

// Enqueue the launches of one schedule on the library stream (what a graph
// captures).  The persistent counter is cleared by a kernel, not a memset.
static int enqueue_schedule(tpz_sw_batch* b, int32_t schedule) {
    const SwPlan& pl = *b->plan;
    const int rows = pl.cfg.rows_per_chunk, W = pl.cfg.waves_per_group;
    SwKernelArgs a = make_args(b);
    const DeviceTasks& dt = pl.packed_tasks;
    a.pair_index = (const int32_t*)dt.d_index.ptr;
    a.entry_rec = (const int4*)dt.d_rec.ptr;
    a.task_cols = (const int32_t*)dt.d_cols.ptr;
    a.task_begin = (const int32_t*)dt.d_begin.ptr;
    a.task_query = (const int32_t*)dt.d_query.ptr;
    // the timeout flag of the staircase waits is read after every run: clear
    // it with a kernel too (a graph replays it)
    if (check(tpz_launch_fill_i32(a.error, 0, 1, stream()), "clear error flag")) return -1;
    if (b->lanes) {
        // the reference's lanes: all pairs in one launch (persistent) or
        // each query's pairs in one (plain)
        LaneArgs la = lane_args(a, b->bias, (int32_t*)b->d_lane_ws.ptr, b->lane_ws_rows);
        if (schedule == TPZ_SCHED_PERSISTENT) {
            la.n_list = b->n_pairs;
            return check(tpz_launch_sw_lane(&la, b->lane_blocks, stream()), "sw_lane");
        }
        for (int32_t q = 0; q + 1 < (int32_t)pl.plain_begin.size(); ++q) {
            const int32_t i0 = pl.plain_begin[q], i1 = pl.plain_begin[q + 1];
            if (i1 == i0) continue;
            la.pairs = (const int32_t*)pl.d_plain_pairs.ptr + i0;
            la.n_list = i1 - i0;
            const int blocks = std::min(b->lane_blocks, (i1 - i0 + kLaneGroupsPerBlock - 1) /
                                                            kLaneGroupsPerBlock);
            if (check(tpz_launch_sw_lane(&la, blocks, stream()), "sw_lane (plain)")) return -1;
        }
        return 0;
    }
    if (schedule == TPZ_SCHED_PERSISTENT) {
        const int n_items = (int)(pl.items_persistent.size() / W);
        if (n_items > 0) {
            a.item_task = (const int32_t*)pl.d_items_persistent.ptr;
            a.n_items = n_items;
            a.task_counter = (int32_t*)b->d_counter.ptr;
            // the counter, the error flag (a run reports only its own timed-out
            // waits) and the keys start at zero in every run, replays included
            if (check(tpz_launch_fill_i32(a.task_counter, 0, 1, stream()), "clear counter") ||
                check(tpz_launch_fill_i32(a.error, 0, 1, stream()), "clear error flag") ||
                check(tpz_launch_fill_i32((int32_t*)a.out_key, 0, 2 * b->n_pairs, stream()),
                      "clear keys"))
                return -1;
            const int grid = (int)std::min<size_t>((size_t)n_items, pl.ws_slots);
            if (check(tpz_launch_sw_forward(&a, grid, rows, W, stream()), "sw_forward (persistent)") ||
                check(tpz_launch_sw_decode(&a, b->n_pairs, stream()), "sw_decode"))
                return -1;
        }
    } else {
        // one launch per query, one wavefront per pair of that query
        PairArgs pa{};
        pa.q_res = a.q_res; pa.q_off = a.q_off; pa.q_len = a.q_len;
        pa.s_res = a.s_res; pa.s_off = a.s_off; pa.s_len = a.s_len;
        pa.pair_query = a.pair_query; pa.pair_subject = a.pair_subject;
        pa.matrix = a.matrix; pa.gap_open = a.gap_open; pa.gap_extend = a.gap_extend;
        pa.workspace = (int32_t*)pl.d_plain_ws.ptr;
        pa.out_score = a.out_score; pa.out_ref = a.out_ref; pa.out_read = a.out_read;
        for (int32_t q = 0; q + 1 < (int32_t)pl.plain_begin.size(); ++q) {
            const int32_t i0 = pl.plain_begin[q], i1 = pl.plain_begin[q + 1];
            if (i1 == i0) continue;
            pa.pairs = (const int32_t*)pl.d_plain_pairs.ptr + i0;
            pa.ws_off = (const int64_t*)pl.d_plain_ws_off.ptr + i0;
            pa.n_list = i1 - i0;
            if (check(tpz_launch_sw_pair(&pa, tpz_sw_pair_rows(b->q_len[q]), stream()),
                      "sw_pair (plain)"))
                return -1;
        }
    }
    return 0;
}

static int capture_schedule(tpz_sw_batch* b, int32_t schedule) {
    SwPlan& pl = *b->plan;
    hipGraph_t g = nullptr;
    if (check(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal), "begin capture"))
        return -1;
    const int rc = enqueue_schedule(b, schedule);
    const hipError_t e = hipStreamEndCapture(stream(), &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return -1; }
    if (check(e, "end capture")) return -1;
    hipGraphExec_t x = nullptr;
    if (check(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "instantiate graph")) {
        (void)hipGraphDestroy(g);
        return -1;
    }
    pl.graph[schedule] = g;
    pl.exec[schedule] = x;
    return 0;
}

int tpz_sw_batch_run(tpz_sw_batch* b, int32_t schedule) {
    if (!b) return fail("null batch");
    if (schedule != TPZ_SCHED_PLAIN && schedule != TPZ_SCHED_PERSISTENT)
        return fail("unknown schedule %d", schedule);
    SwPlan& pl = *b->plan;
    if (pl.cfg.use_graph && !pl.exec[schedule] && capture_schedule(b, schedule)) return -1;
    if (check(hipEventRecord(b->ev[0], stream()), "event")) return -1;
    if (pl.cfg.use_graph) {
        if (check(hipGraphLaunch(pl.exec[schedule], stream()), "graph launch")) return -1;
    } else if (enqueue_schedule(b, schedule)) {
        return -1;
    }
    if (check(hipEventRecord(b->ev[1], stream()), "event")) return -1;
    if (check(hipStreamSynchronize(stream()), "sw_forward")) return -1;
    float ms = 0;
    if (check(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]), "elapsed")) return -1;
    b->ms = ms;
    int32_t err = 0;
    if (check(hipMemcpy(&err, b->d_error.ptr, 4, hipMemcpyDeviceToHost), "error flag")) return -1;
    if (err) return fail("sw_forward: a cross-wavefront wait timed out");
    return 0;
}

int tpz_sw_batch_get_config(const tpz_sw_batch* b, tpz_sw_config* cfg) {
    if (!b || !cfg) return fail("null argument");
    *cfg = b->plan->cfg;
    return 0;
}

int tpz_sw_batch_set_config(tpz_sw_batch* b, const tpz_sw_config* cfg) {
    if (!b || !cfg) return fail("null argument");
    return build_plan(b, *cfg);
}

int tpz_sw_batch_autotune(tpz_sw_batch* b, tpz_sw_config* best, double* best_ms,
                          int32_t* n_tried) {
    if (!b) return fail("null batch");
    const tpz_sw_config start = b->plan->cfg;
    std::vector<tpz_sw_config> cands;
    for (int rows : {32, 16}) {
        const int max_groups = std::max(1, tpz_sw_forward_groups_per_cu(rows, 4));
        for (int groups = 1; groups <= max_groups; ++groups)
            for (int nq : {4, 2, 1})
                for (double bf : {0.25, 0.5, 1.0, 1e30}) {
                    if ((nq != 4 || groups != max_groups) && bf != 0.5) continue;
                    tpz_sw_config c = start;
                    c.rows_per_chunk = rows; c.groups_per_cu = groups;
                    c.queries_per_task = nq; c.big_fraction = bf; c.use_graph = 0;
                    c.waves_per_group = 4;
                    cands.push_back(c);
                }
    }
    // eight wavefronts per workgroup (one workgroup per CU), 32 rows
    if (tpz_sw_forward_groups_per_cu(32, 8) >= 1)
        for (double bf : {0.5, 1.0}) {
            tpz_sw_config c = start;
            c.rows_per_chunk = 32; c.groups_per_cu = 1; c.queries_per_task = 4;
            c.big_fraction = bf; c.use_graph = 0; c.waves_per_group = 8;
            cands.push_back(c);
        }
    const size_t n = (size_t)b->n_pairs;
    std::vector<int32_t> ref_sc(n), ref_re(n), ref_rd(n), sc(n), re(n), rd(n);
    double best_t = 1e300;
    tpz_sw_config best_c = start;
    for (size_t k = 0; k < cands.size(); ++k) {
        if (build_plan(b, cands[k]) || tpz_sw_batch_run(b, TPZ_SCHED_PERSISTENT)) return -1;
        double t = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            if (tpz_sw_batch_run(b, TPZ_SCHED_PERSISTENT)) return -1;
            t = std::min(t, b->ms);
        }
        if (tpz_sw_batch_results(b, k ? sc.data() : ref_sc.data(), k ? re.data() : ref_re.data(),
                                 k ? rd.data() : ref_rd.data()))
            return -1;
        if (k && (sc != ref_sc || re != ref_re || rd != ref_rd))
            return fail("autotune: configuration %zu (rows %d, groups %d, queries %d) changed "
                        "the results", k, cands[k].rows_per_chunk, cands[k].groups_per_cu,
                        cands[k].queries_per_task);
        if (t < best_t) { best_t = t; best_c = cands[k]; }
    }
    best_c.use_graph = start.use_graph;
    if (build_plan(b, best_c)) return -1;
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    if (n_tried) *n_tried = (int32_t)cands.size();
    return 0;
}

int tpz_sw_batch_results(const tpz_sw_batch* b, int32_t* score, int32_t* ref_end,
                         int32_t* read_end) {
    if (!b) return fail("null batch");
    const size_t n = (size_t)b->n_pairs * 4;
    if (score && n && check(hipMemcpy(score, b->d_score.ptr, n, hipMemcpyDeviceToHost), "results")) return -1;
    if (ref_end && n && check(hipMemcpy(ref_end, b->d_ref.ptr, n, hipMemcpyDeviceToHost), "results")) return -1;
    if (read_end && n && check(hipMemcpy(read_end, b->d_read.ptr, n, hipMemcpyDeviceToHost), "results")) return -1;
    return 0;
}

int tpz_sw_batch_kernel_ms(const tpz_sw_batch* b, double* ms) {
    if (!b) return fail("null batch");
    if (ms) *ms = b->ms;
    return 0;
}

void tpz_sw_batch_destroy(tpz_sw_batch* b) {
    if (!b) return;
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    delete b;
}


// ---------------------------------------------------------------------------
// gapless extension
// ---------------------------------------------------------------------------
}  // extern "C"

namespace tpz {

// lexicographical.h:19-33 (aa_table2) and :60-74 (upper_aa_only): the
// internal alphabet stores amino acid k of "ACMLJIVTSKRQZEDBNHYFWXP*G" as
// 'A'+k (either case); every other byte maps to matrix row 23.  Query bytes
// other than 'A'..'Y', '*' and '\n' score as 'V'.
static void alphabet_maps(int8_t* tmap, int8_t* qmap) {
    static const char internal_order[] = "ACMLJIVTSKRQZEDBNHYFWXP*G";
    static const char matrix_order[] = "ARNDCQEGHILKMFPSTWYVBJZX*";
    for (int c = 0; c < 128; ++c) tmap[c] = 23;
    for (int k = 0; k < 25; ++k) {
        const int8_t idx = (int8_t)(strchr(matrix_order, internal_order[k]) - matrix_order);
        tmap['A' + k] = idx;
        tmap['a' + k] = idx;
    }
    for (int c = 0; c < 128; ++c) {
        const bool keep = (c >= 'A' && c <= 'Y') || c == '*' || c == '\n';
        qmap[c] = keep ? tmap[c] : tmap['V'];
    }
}

// Matrix rows of a text byte and of a query byte (bytes >= 128 use their low
// seven bits, as the reference's 128-entry tables would need).
struct AlphabetMaps {
    int8_t tmap[128], qmap[128];
    AlphabetMaps() { alphabet_maps(tmap, qmap); }
};
static const AlphabetMaps& maps() {
    static const AlphabetMaps m;
    return m;
}
static int text_row(uint8_t c) { return maps().tmap[c & 127]; }
void alphabet_rows(int8_t* text_rows, int8_t* query_rows) {
    memcpy(text_rows, maps().tmap, 128);
    memcpy(query_rows, maps().qmap, 128);
}
static int query_row(uint8_t c) { return maps().qmap[c & 127]; }

// Codes by letter frequency: the gapless kernel's table lookups hit LDS bank
// (code >> 1), so the 16 most frequent letters get the even codes 0..30, one
// bank each, and rarer letters the odd codes 29, 27, ... (sharing banks with
// the less frequent of the first sixteen).  With codes in byte order the
// twenty amino acids fell into ten banks.
static int text_codes(const uint8_t* text, uint64_t n, TextCodes& tc) {
    // byte counts, per host thread over 1 MB pieces, then summed
    const int threads = n >= ((uint64_t)1 << 24) ? host_threads() : 1;
    const int64_t pieces = (int64_t)((n + (1 << 20) - 1) >> 20);
    std::vector<std::array<uint64_t, 256>> part((size_t)std::max(1, threads));
    for (auto& c : part) c.fill(0);
    std::atomic<int> slot{0};
    parallel_for(pieces, threads, [&](int64_t p0, int64_t p1) {
        std::array<uint64_t, 256>& c = part[(size_t)slot++];
        const uint64_t e = std::min<uint64_t>(n, (uint64_t)p1 << 20);
        for (uint64_t i = (uint64_t)p0 << 20; i < e; ++i) ++c[text[i]];
    });
    uint64_t count[256] = {0};
    for (const auto& c : part)
        for (int b = 0; b < 256; ++b) count[b] += c[b];
    int letters[256], n_letters = 0;
    for (int c = 0; c < 256; ++c)
        if (count[c] && c != '\n') letters[n_letters++] = c;
    if (n_letters > kTextTerm)
        return fail("the text holds more than %d distinct letters", kTextTerm);
    std::stable_sort(letters, letters + n_letters,
                     [&](int a, int b) { return count[a] > count[b]; });
    memset(tc.code, 0xff, sizeof tc.code);
    memset(tc.byte_of, 0, sizeof tc.byte_of);
    memset(tc.used, 0, sizeof tc.used);
    for (int r = 0; r < n_letters; ++r) {
        const int code = r < 16 ? 2 * r : 2 * (30 - r) + 1;     // ranks 16..30 -> 29, 27, .., 1
        tc.byte_of[code] = (uint8_t)letters[r];
        tc.code[letters[r]] = (uint8_t)code;
        tc.used[code] = true;
    }
    tc.code['\n'] = kTextTerm;
    tc.byte_of[kTextTerm] = '\n';
    return 0;
}

static uint8_t query_code(const TextCodes& tc, uint8_t c) {
    return tc.code[c] != 0xff ? tc.code[c] : (uint8_t)(kForeignCode + query_row(c));
}

int gapless_query_codes(const TextCodes& tc, int32_t n_queries, const char* q_chars,
                        const int64_t* q_off, const int32_t* q_len, QueryCodes& out) {
    out.fwd.clear(); out.rev.clear();
    out.fwd_off.assign(n_queries, 0);
    out.rev_off.assign(n_queries, 0);
    for (int32_t k = 0; k < n_queries; ++k) {
        if (q_len[k] < 0) return fail("query %d has negative length", k);
        const uint8_t* q = reinterpret_cast<const uint8_t*>(q_chars + q_off[k]);
        const int32_t m = q_len[k];
        out.fwd_off[k] = (int64_t)out.fwd.size();
        for (int32_t j = 0; j < m; ++j) out.fwd.push_back((uint8_t)(2 * query_code(tc, q[j])));
        out.fwd.push_back(2 * kQueryTerm);
        // reversed: q[m-1] .. q[1], then q[0]'s slot and one more terminator
        out.rev_off[k] = (int64_t)out.rev.size();
        for (int32_t j = m - 1; j >= 1; --j) out.rev.push_back((uint8_t)(2 * query_code(tc, q[j])));
        out.rev.push_back(2 * kQueryTerm);
        out.rev.push_back(2 * kQueryTerm);
    }
    out.fwd.insert(out.fwd.end(), kTextPad, 2 * kQueryTerm);
    out.rev.insert(out.rev.end(), kTextPad, 2 * kQueryTerm);
    return 0;
}

void gapless_table(const TextCodes& tc, const int8_t* matrix, int16_t* table) {
    for (int t = 0; t < kTableRows; ++t) {
        for (int q = 0; q < kTableCols; ++q) {
            int w = kTermScore;
            if (t < kTextTerm && tc.used[t] && q != kQueryTerm) {
                const int tr = text_row(tc.byte_of[t]);
                int qr = -1;
                if ((q < kTextTerm && tc.used[q]) || q == kTextTerm) qr = query_row(tc.byte_of[q]);
                else if (q >= kForeignCode && q < kForeignCode + 25) qr = q - kForeignCode;
                if (qr >= 0) w = matrix[tr * 25 + qr];
            }
            table[t * kTableCols + q] = (int16_t)w;
        }
    }
}

}  // namespace tpz

extern "C" {

int tpz_text_create(tpz_text** out, const char* text, uint64_t t_size, const int64_t* sa) {
    if (!out) return fail("tpz_text_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    if (!text && t_size) return fail("tpz_text_create: null text");
    auto t = new tpz_text();
    t->t_size = t_size;
    t->has_sa = sa != nullptr;
    const uint8_t* T = reinterpret_cast<const uint8_t*>(text);
    if (text_codes(T, t_size, t->codes)) { delete t; return -1; }
    // the code copy followed by terminators, and the head window: the left
    // scan stops before T[0] (i < idx), so there T[0] is a terminator
    const uint8_t* code = t->codes.code;
    int rc = t->d_fwd.alloc(t_size + kTextPad);
    if (!rc)
        rc |= stage_to_device(t->d_fwd.ptr, t_size + kTextPad, [&](uint8_t* out, size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) out[i - b] = i < t_size ? code[T[i]] : (uint8_t)kTextTerm;
        });
    std::vector<uint8_t> head(kHeadBytes, kTextTerm);
    for (uint64_t i = 1; i < t_size && i < (uint64_t)kWindow; ++i) head[kWindow + i] = code[T[i]];
    rc |= t->d_head.upload(head.data(), head.size());
    if (!rc && sa) rc |= t->d_sa.upload_large(sa, t_size * 8);
    if (!rc) rc |= check(hipStreamSynchronize(stream()), "text upload");
    if (rc) { delete t; return -1; }
    *out = t;
    return 0;
}

void tpz_text_destroy(tpz_text* t) { delete t; }

struct tpz_gapless_batch {
    const tpz_text* text = nullptr;
    int32_t n_queries = 0;
    int64_t n_seeds = 0;
    DeviceBuffer d_qf, d_qf_off, d_qr, d_qr_off, d_q_len, d_seeds, d_table, d_score, d_len;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms = 0;
};

int tpz_gapless_batch_create(tpz_gapless_batch** out, const tpz_text* text,
                             int32_t n_queries, const char* q_chars, const int64_t* q_off,
                             const int32_t* q_len, int64_t n_seeds, const tpz_seed* seeds,
                             const int8_t* matrix) {
    if (!out) return fail("tpz_gapless_batch_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    if (!text) return fail("null text");
    if (n_queries < 0 || n_seeds < 0) return fail("negative counts");
    if (!matrix) return fail("null matrix");
    QueryCodes qc;
    if (gapless_query_codes(text->codes, n_queries, q_chars, q_off, q_len, qc)) return -1;
    for (int64_t i = 0; i < n_seeds; ++i) {
        if (seeds[i].query < 0 || seeds[i].query >= n_queries)
            return fail("seed %lld names query %d out of range", (long long)i, seeds[i].query);
        if (text->has_sa && (seeds[i].position < 0 || (uint64_t)seeds[i].position >= text->t_size))
            return fail("seed %lld names suffix-array slot %lld outside [0, %llu)", (long long)i,
                        (long long)seeds[i].position, (unsigned long long)text->t_size);
    }
    std::vector<int16_t> table(kTableRows * kTableCols);
    gapless_table(text->codes, matrix, table.data());
    auto b = new tpz_gapless_batch();
    b->text = text;
    b->n_queries = n_queries;
    b->n_seeds = n_seeds;
    int rc = 0;
    rc |= b->d_qf.upload(qc.fwd.data(), qc.fwd.size());
    rc |= b->d_qr.upload(qc.rev.data(), qc.rev.size());
    rc |= b->d_qf_off.upload(qc.fwd_off.data(), qc.fwd_off.size() * 8);
    rc |= b->d_qr_off.upload(qc.rev_off.data(), qc.rev_off.size() * 8);
    rc |= b->d_q_len.upload(q_len, (size_t)n_queries * 4);
    rc |= b->d_seeds.upload(seeds, (size_t)n_seeds * sizeof(tpz_seed));
    rc |= b->d_table.upload(table.data(), table.size() * 2);
    rc |= b->d_score.alloc((size_t)n_seeds * 4 + 4);
    rc |= b->d_len.alloc((size_t)n_seeds * 4 + 4);
    for (auto& e : b->ev) rc |= check(hipEventCreate(&e), "hipEventCreate");
    rc |= check(hipStreamSynchronize(stream()), "gapless upload");
    if (rc) { tpz_gapless_batch_destroy(b); return -1; }
    *out = b;
    return 0;
}

int tpz_gapless_batch_run(tpz_gapless_batch* b) {
    if (!b) return fail("null batch");
    GaplessArgs a{};
    a.t_fwd = (const uint8_t*)b->text->d_fwd.ptr;
    a.t_head = (const uint8_t*)b->text->d_head.ptr;
    a.t_size = (int64_t)b->text->t_size;
    a.sa = b->text->has_sa ? (const int64_t*)b->text->d_sa.ptr : nullptr;
    a.q_fwd = (const uint8_t*)b->d_qf.ptr;
    a.qf_off = (const int64_t*)b->d_qf_off.ptr;
    a.q_rev = (const uint8_t*)b->d_qr.ptr;
    a.qr_off = (const int64_t*)b->d_qr_off.ptr;
    a.q_len = (const int32_t*)b->d_q_len.ptr;
    a.n_queries = b->n_queries;
    a.seeds = (const SeedRec*)b->d_seeds.ptr;
    a.n_seeds = b->n_seeds;
    a.table = (const int16_t*)b->d_table.ptr;
    a.out_score = (int32_t*)b->d_score.ptr;
    a.out_length = (int32_t*)b->d_len.ptr;
    const int64_t blocks_needed = (b->n_seeds + 255) / 256;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(blocks_needed, (int64_t)num_cu() * 32));
    if (check(hipEventRecord(b->ev[0], stream()), "event")) return -1;
    if (b->n_seeds > 0 && check(tpz_launch_gapless(&a, grid, stream()), "gapless")) return -1;
    if (check(hipEventRecord(b->ev[1], stream()), "event")) return -1;
    if (check(hipStreamSynchronize(stream()), "gapless")) return -1;
    float ms = 0;
    if (check(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]), "elapsed")) return -1;
    b->ms = ms;
    return 0;
}

int tpz_gapless_batch_results(const tpz_gapless_batch* b, int32_t* score, int32_t* length) {
    if (!b) return fail("null batch");
    const size_t n = (size_t)b->n_seeds * 4;
    if (score && n && check(hipMemcpy(score, b->d_score.ptr, n, hipMemcpyDeviceToHost), "results")) return -1;
    if (length && n && check(hipMemcpy(length, b->d_len.ptr, n, hipMemcpyDeviceToHost), "results")) return -1;
    return 0;
}

int tpz_gapless_batch_kernel_ms(const tpz_gapless_batch* b, double* ms) {
    if (!b) return fail("null batch");
    if (ms) *ms = b->ms;
    return 0;
}

void tpz_gapless_batch_destroy(tpz_gapless_batch* b) {
    if (!b) return;
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    delete b;
}

// ---------------------------------------------------------------------------
// traceback statistics
// ---------------------------------------------------------------------------
}  // extern "C"

namespace tpz {

char* BlockList::take(size_t bytes, void* (*grow)(size_t, void*), void* ctx) {
    bytes = (bytes + 255) & ~(size_t)255;
    while (cur < blocks.size() && off + bytes > blocks[cur].second) {
        ++cur;
        off = 0;
    }
    if (cur == blocks.size()) {
        const size_t last = blocks.empty() ? 0 : blocks.back().second;
        const size_t size = std::max(std::max(bytes, 2 * last), (size_t)1 << 20);
        char* p = (char*)grow(size, ctx);
        if (!p) return nullptr;
        blocks.emplace_back(p, size);
        off = 0;
    }
    char* p = blocks[cur].first + off;
    off += bytes;
    return p;
}

int stats_check(const StatsInput& in) {
    if (in.n_queries < 0 || in.n_subjects < 0 || in.n_hits < 0) return fail("negative counts");
    if (!gaps_valid(in.gap_open, in.gap_extend))
        return fail("gap penalties must lie in [0, 255] (got %d/%d)", in.gap_open, in.gap_extend);
    if (!in.matrix) return fail("null matrix");
    for (int32_t h = 0; h < in.n_hits; ++h) {
        const int32_t qi = in.hit_query[h], si = in.hit_subject[h];
        if (qi < 0 || qi >= in.n_queries || si < 0 || si >= in.n_subjects)
            return fail("hit %d references query %d / subject %d out of range", h, qi, si);
        if (in.read_end1[h] < 0 || in.read_end1[h] >= in.q_len[qi] || in.ref_end1[h] < 0 ||
            in.ref_end1[h] >= in.s_len[si] || in.score1[h] <= 0 || in.score1[h] >= 32767)
            return fail("hit %d: end positions (%d, %d) / score %d outside the query, "
                        "subject or (0, 32767)", h, in.read_end1[h], in.ref_end1[h], in.score1[h]);
    }
    for (int32_t k = 0; k < in.n_queries; ++k)
        for (int32_t j = 0; j < in.q_len[k]; ++j)
            if (in.q_res[in.q_off[k] + j] < 0 || in.q_res[in.q_off[k] + j] > 24)
                return fail("query %d residue out of range", k);
    for (int32_t k = 0; k < in.n_subjects; ++k)
        for (int32_t j = 0; j < in.s_len[k]; ++j)
            if (in.s_res[in.s_off[k] + j] < 0 || in.s_res[in.s_off[k] + j] > 24)
                return fail("subject %d residue out of range", k);
    return 0;
}

namespace {
// host block -> device block of the same size, on the run's stream
template <typename T>
T* put(RunMem& m, const T* src, size_t count, size_t pad = 0) {
    const size_t bytes = count * sizeof(T) + pad;
    char* h = m.host(bytes);
    char* d = (char*)m.dev(bytes);
    if (!h || !d) return nullptr;
    if (count) memcpy(h, src, count * sizeof(T));
    if (pad) memset(h + count * sizeof(T), 0, pad);
    if (check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, m.stream()), "upload"))
        return nullptr;
    return (T*)d;
}
// device -> pinned host block, read after the next wait
int32_t* get(RunMem& m, const void* d, int32_t n) {
    int32_t* h = (int32_t*)m.host((size_t)n * 4 + 4);
    if (!h) return nullptr;
    if (check(hipMemcpyAsync(h, d, (size_t)n * 4, hipMemcpyDeviceToHost, m.stream()), "readback"))
        return nullptr;
    return h;
}
}  // namespace

// The reverse pass gives the start coordinates (ssw.c:900-941), the band
// search the band half-width (ssw.c:570-660), the trace the counts
// (ssw.c:663-739); each stage's sizes come from the one before, so the
// host waits three times.  All memory comes from m and is reused after the
// next reset, so a caller that keeps m allocates nothing per run.
int stats_compute(const StatsInput& in, RunMem& m, StatsOutput& out) {
    const int32_t n = in.n_hits;
    out.rc.assign(n, 0); out.read_begin.assign(n, 0); out.ref_begin.assign(n, 0);
    out.len.assign(n, 0); out.gapo.assign(n, 0); out.mis.assign(n, 0); out.ident.assign(n, 0);
    if (n == 0) return 0;
    m.reset();
    int32_t minw = 0;
    for (int i = 0; i < 625; ++i) minw = std::min<int32_t>(minw, in.matrix[i]);
    int64_t qtot = 0, stot = 0;
    for (int32_t k = 0; k < in.n_queries; ++k) qtot = std::max<int64_t>(qtot, in.q_off[k] + in.q_len[k]);
    for (int32_t k = 0; k < in.n_subjects; ++k) stot = std::max<int64_t>(stot, in.s_off[k] + in.s_len[k]);
    // reverse-pass workspace: bottom-row H and F between passes and the
    // column maxima, one word per subject column each
    std::vector<int64_t> ws_off(n);
    int64_t ws = 0;
    for (int32_t h = 0; h < n; ++h) {
        ws_off[h] = ws;
        ws += 3 * ((int64_t)in.ref_end1[h] + 1);
    }
    ReverseArgs r{};
    r.q_res = put(m, in.q_res, (size_t)qtot, 16);
    r.q_off = put(m, in.q_off, (size_t)in.n_queries);
    r.s_res = put(m, in.s_res, (size_t)stot, 16);
    r.s_off = put(m, in.s_off, (size_t)in.n_subjects);
    r.pair_query = put(m, in.hit_query, (size_t)n);
    r.pair_subject = put(m, in.hit_subject, (size_t)n);
    r.read_end = put(m, in.read_end1, (size_t)n);
    r.ref_end = put(m, in.ref_end1, (size_t)n);
    r.score = put(m, in.score1, (size_t)n);
    r.matrix = put(m, in.matrix, 625);
    r.ws_off = put(m, ws_off.data(), (size_t)n);
    r.workspace = (int32_t*)m.dev((size_t)ws * 4 + 16);
    r.out_read_begin = (int32_t*)m.dev((size_t)n * 4 + 4);
    r.out_ref_begin = (int32_t*)m.dev((size_t)n * 4 + 4);
    if (!r.q_res || !r.q_off || !r.s_res || !r.s_off || !r.pair_query || !r.pair_subject ||
        !r.read_end || !r.ref_end || !r.score || !r.matrix || !r.ws_off || !r.workspace ||
        !r.out_read_begin || !r.out_ref_begin)
        return -1;
    r.n_pairs = n;
    r.gap_open = in.gap_open; r.gap_extend = in.gap_extend; r.bias = -minw;
    if (!textbook_gaps(in.gap_open, in.gap_extend)) {
        // the reference's lanes over the reversed prefixes, all hits at once
        LaneArgs la{};
        la.q_res = r.q_res; la.q_off = r.q_off; la.s_res = r.s_res; la.s_off = r.s_off;
        la.pair_query = r.pair_query; la.pair_subject = r.pair_subject;
        la.n_list = n;
        la.matrix = r.matrix;
        la.gap_open = in.gap_open; la.gap_extend = in.gap_extend; la.bias = -minw;
        la.read_end = r.read_end; la.ref_end = r.ref_end; la.score = r.score;
        int32_t max_prefix = 0;
        for (int32_t h = 0; h < n; ++h) max_prefix = std::max(max_prefix, in.read_end1[h] + 1);
        la.ws_rows = lane_rows(max_prefix);
        const int blocks = (int)std::min<int64_t>((n + kLaneGroupsPerBlock - 1) / kLaneGroupsPerBlock,
                                                  (int64_t)std::max(num_cu(), 1) * 16);
        la.workspace = (int32_t*)m.dev((size_t)blocks * kLaneGroupsPerBlock * 3 *
                                       (size_t)la.ws_rows * 4 + 64);
        la.out_ref = r.out_ref_begin;
        la.out_read = r.out_read_begin;
        if (!la.workspace || check(tpz_launch_sw_lane(&la, blocks, m.stream()), "sw_lane (reverse)"))
            return -1;
    } else {
        // one launch per rows-per-lane class of the hits' lane-padded prefixes
        std::vector<int32_t> cls[3];
        for (int32_t h = 0; h < n; ++h) {
            const int32_t L = in.read_end1[h] + 1;
            const int32_t P = -minw + in.score1[h] < 255 ? 16 : 8;
            const int rpl = reverse_rows_per_lane((L + P - 1) / P * P);
            cls[rpl == 8 ? 0 : rpl == 16 ? 1 : 2].push_back(h);
        }
        for (int c = 0; c < 3; ++c) {
            if (cls[c].empty()) continue;
            r.hits = put(m, cls[c].data(), cls[c].size());
            r.n_list = (int32_t)cls[c].size();
            if (!r.hits || check(tpz_launch_sw_reverse(&r, 8 << c, m.stream()), "sw_reverse"))
                return -1;
        }
    }
    const int32_t* hrb = get(m, r.out_ref_begin, n);
    const int32_t* hqb = get(m, r.out_read_begin, n);
    if (!hrb || !hqb || m.wait()) return -1;
    out.ref_begin.assign(hrb, hrb + n);
    out.read_begin.assign(hqb, hqb + n);
    const std::vector<int32_t>& rb = out.ref_begin;
    const std::vector<int32_t>& qb = out.read_begin;
    for (int32_t h = 0; h < n; ++h)
        if (rb[h] < 0 || qb[h] < 0 || qb[h] > in.read_end1[h])
            return fail("hit %d: reverse pass found no start (score1 %d not reached)", h, in.score1[h]);
    BandArgs a{};
    a.q_res = r.q_res; a.q_off = r.q_off; a.s_res = r.s_res; a.s_off = r.s_off;
    a.pair_query = r.pair_query; a.pair_subject = r.pair_subject;
    a.read_begin = r.out_read_begin;
    a.read_end = r.read_end;
    a.ref_begin = r.out_ref_begin;
    a.ref_end = r.ref_end;
    a.score = r.score;
    a.n_pairs = n;
    a.matrix = r.matrix;
    a.gap_open = in.gap_open; a.gap_extend = in.gap_extend;
    a.out_band = (int32_t*)m.dev((size_t)n * 4 + 4);
    if (!a.out_band) return -1;
    // band search with a wavefront per hit; a band wider than the LDS rows
    // (status -2) goes again through the one-thread-per-hit kernel, whose
    // rows live in the workspace
    a.hits = nullptr;
    a.n_list = n;
    if (check(tpz_launch_band_search(&a, 1, m.stream()), "band_search")) return -1;
    const int32_t* band = get(m, a.out_band, n);
    if (!band || m.wait()) return -1;
    std::vector<int32_t> wide;
    for (int32_t h = 0; h < n; ++h)
        if (band[h] == -2) wide.push_back(h);
    // workspace band rows for those hits only, each sized by its own
    // rectangle (the trace's wide hits are among them: their final band is
    // at least the one that did not fit the LDS rows)
    std::vector<int64_t> bo(n, 0);
    int64_t bw = 0;
    for (int32_t h : wide) {
        bo[h] = bw;
        bw += 3 * (int64_t)band_row_words(in.ref_end1[h] - rb[h] + 1, in.read_end1[h] - qb[h] + 1);
    }
    if (!wide.empty()) {
        a.ws_off = put(m, bo.data(), (size_t)n);
        a.workspace = (int32_t*)m.dev((size_t)bw * 4 + 16);
        a.hits = put(m, wide.data(), wide.size());
        a.n_list = (int32_t)wide.size();
        if (!a.ws_off || !a.workspace || !a.hits ||
            check(tpz_launch_band_search(&a, 0, m.stream()), "band_search"))
            return -1;
        band = get(m, a.out_band, n);
        if (!band || m.wait()) return -1;
    }
    // direction buffers: the reference's size, 1024 bytes grown to the next
    // power of two above each pass's (2w+1) * readLen * 3 (ssw.c:555, 591-597)
    std::vector<int64_t> dofs(n), dsize(n);
    int64_t dtot = 0;
    for (int32_t h = 0; h < n; ++h) {
        dofs[h] = dtot;
        int64_t s2 = 1024;
        if (band[h] >= 0) {
            const int64_t ncell = (int64_t)(2 * band[h] + 1) * (in.read_end1[h] - qb[h] + 1) * 3;
            while (ncell >= s2) {
                int64_t v = s2;           // ++s2 then round up to a power of two
                v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
                s2 = v + 1;
            }
        }
        dsize[h] = band[h] >= 0 ? s2 : 0;
        dtot += dsize[h];
    }
    a.dirs = (int8_t*)m.dev((size_t)dtot + 16);
    a.dir_off = put(m, dofs.data(), (size_t)n);
    a.dir_size = put(m, dsize.data(), (size_t)n);
    int32_t* d_res = (int32_t*)m.dev(5 * ((size_t)n * 4 + 4));
    if (!a.dirs || !a.dir_off || !a.dir_size || !d_res ||
        check(hipMemsetAsync(a.dirs, 0, (size_t)dtot + 16, m.stream()), "memset"))
        return -1;
    a.out_rc = d_res;
    a.out_length = d_res + (n + 1);
    a.out_gapopen = d_res + 2 * (n + 1);
    a.out_mismatch = d_res + 3 * (n + 1);
    a.out_identical = d_res + 4 * (n + 1);
    // trace: a wavefront per hit whose final band fits the LDS rows
    std::vector<int32_t> narrow;
    wide.clear();
    for (int32_t h = 0; h < n; ++h)
        (band[h] >= 0 && 2 * band[h] + 3 > kBandSlots ? wide : narrow).push_back(h);
    for (int pass = 0; pass < 2; ++pass) {
        const std::vector<int32_t>& list = pass == 0 ? narrow : wide;
        if (list.empty()) continue;
        if (pass == 1 && !a.workspace) {
            a.ws_off = put(m, bo.data(), (size_t)n);
            a.workspace = (int32_t*)m.dev((size_t)bw * 4 + 16);
            if (!a.ws_off || !a.workspace) return -1;
        }
        a.hits = put(m, list.data(), list.size());
        a.n_list = (int32_t)list.size();
        if (!a.hits || check(tpz_launch_band_trace(&a, pass == 0, m.stream()), "band_trace"))
            return -1;
    }
    const int32_t* res = get(m, d_res, 5 * (n + 1));
    if (!res || m.wait()) return -1;
    out.rc.assign(res, res + n);
    out.len.assign(res + (n + 1), res + (n + 1) + n);
    out.gapo.assign(res + 2 * (n + 1), res + 2 * (n + 1) + n);
    out.mis.assign(res + 3 * (n + 1), res + 3 * (n + 1) + n);
    out.ident.assign(res + 4 * (n + 1), res + 4 * (n + 1) + n);
    return 0;
}

double stats_identity(int32_t ident, int32_t length) {
    // ssw.c:741: identical / (double) length; -1.0 / 0.0 for the size guard /
    // trace failure (identity is never assigned there)
    return ident == -1 ? -1.0 : ident == -2 ? 0.0 : ident / (double)length;
}

}  // namespace tpz

namespace {
// Memory of a tpz_stats_batch: its own device and pinned blocks, freed
// with the batch; waits synchronise the library stream.
struct BatchMem final : tpz::RunMem {
    tpz::BlockList d, h;
    ~BatchMem() override {
        for (auto& b : d.blocks) (void)hipFree(b.first);
        for (auto& b : h.blocks) (void)hipHostFree(b.first);
    }
    void reset() override { d.reset(); h.reset(); }
    void* dev(size_t bytes) override {
        return d.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            return tpz::check(hipMalloc(&p, n), "hipMalloc") ? nullptr : p;
        }, nullptr);
    }
    char* host(size_t bytes) override {
        return h.take(bytes, [](size_t n, void*) -> void* {
            void* p = nullptr;
            return tpz::check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc") ? nullptr : p;
        }, nullptr);
    }
    hipStream_t stream() override { return tpz::stream(); }
    int wait() override { return tpz::check(hipStreamSynchronize(tpz::stream()), "stats"); }
};
}  // namespace

struct tpz_stats_batch {
    // inputs, copied at create
    std::vector<int8_t> q_res, s_res, matrix;
    std::vector<int64_t> q_off, s_off;
    std::vector<int32_t> q_len, s_len, hq, hs, read_end, ref_end, score;
    tpz::StatsInput in{};
    BatchMem mem;
    tpz::StatsOutput out;
    bool ran = false;
};

extern "C" {

int tpz_stats_batch_create(tpz_stats_batch** out, int32_t n_queries, const int8_t* q_res,
                           const int64_t* q_off, const int32_t* q_len, int32_t n_subjects,
                           const int8_t* s_res, const int64_t* s_off, const int32_t* s_len,
                           int32_t n_hits, const int32_t* hit_query, const int32_t* hit_subject,
                           const int32_t* read_end1, const int32_t* ref_end1,
                           const int32_t* score1, const int8_t* matrix, int32_t gap_open,
                           int32_t gap_extend) {
    if (!out) return fail("tpz_stats_batch_create: null output pointer");
    *out = nullptr;
    if (ensure_init()) return -1;
    StatsInput in{n_queries, q_res, q_off, q_len, n_subjects, s_res, s_off, s_len, n_hits,
                  hit_query, hit_subject, read_end1, ref_end1, score1, matrix, gap_open,
                  gap_extend};
    if (stats_check(in)) return -1;
    auto b = new tpz_stats_batch();
    int64_t qtot = 0, stot = 0;
    for (int32_t k = 0; k < n_queries; ++k) qtot = std::max<int64_t>(qtot, q_off[k] + q_len[k]);
    for (int32_t k = 0; k < n_subjects; ++k) stot = std::max<int64_t>(stot, s_off[k] + s_len[k]);
    b->q_res.assign(q_res, q_res + qtot);
    b->s_res.assign(s_res, s_res + stot);
    b->matrix.assign(matrix, matrix + 625);
    b->q_off.assign(q_off, q_off + n_queries);
    b->s_off.assign(s_off, s_off + n_subjects);
    b->q_len.assign(q_len, q_len + n_queries);
    b->s_len.assign(s_len, s_len + n_subjects);
    b->hq.assign(hit_query, hit_query + n_hits);
    b->hs.assign(hit_subject, hit_subject + n_hits);
    b->read_end.assign(read_end1, read_end1 + n_hits);
    b->ref_end.assign(ref_end1, ref_end1 + n_hits);
    b->score.assign(score1, score1 + n_hits);
    b->in = StatsInput{n_queries, b->q_res.data(), b->q_off.data(), b->q_len.data(), n_subjects,
                       b->s_res.data(), b->s_off.data(), b->s_len.data(), n_hits, b->hq.data(),
                       b->hs.data(), b->read_end.data(), b->ref_end.data(), b->score.data(),
                       b->matrix.data(), gap_open, gap_extend};
    *out = b;
    return 0;
}

int tpz_stats_batch_run(tpz_stats_batch* b) {
    if (!b) return fail("null batch");
    b->ran = false;
    if (stats_compute(b->in, b->mem, b->out)) return -1;
    b->ran = true;
    return 0;
}

int tpz_stats_batch_results(const tpz_stats_batch* b, int32_t* rc, int32_t* read_begin1,
                            int32_t* ref_begin1, int32_t* length, int32_t* gapopen,
                            int32_t* mismatch, double* identity) {
    if (!b) return fail("null batch");
    if (!b->ran && b->in.n_hits > 0) return fail("tpz_stats_batch_results before run");
    const StatsOutput& o = b->out;
    for (int32_t h = 0; h < b->in.n_hits; ++h) {
        if (rc) rc[h] = o.rc[h];
        if (read_begin1) read_begin1[h] = o.read_begin[h];
        if (ref_begin1) ref_begin1[h] = o.ref_begin[h];
        if (length) length[h] = o.len[h];
        if (gapopen) gapopen[h] = o.gapo[h];
        if (mismatch) mismatch[h] = o.mis[h];
        if (identity) identity[h] = stats_identity(o.ident[h], o.len[h]);
    }
    return 0;
}

void tpz_stats_batch_destroy(tpz_stats_batch* b) { delete b; }

}  // extern "C"
