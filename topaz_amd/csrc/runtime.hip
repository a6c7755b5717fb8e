// Host runtime of libtopaz_amd: device selection, error reporting, and the
// batched Smith-Waterman driver behind the C ABI of include/topaz_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/topaz_amd.h"
#include "runtime.h"
#include "sw_kernels.h"

namespace tpz {

// ---------------------------------------------------------------------------
// errors and device
// ---------------------------------------------------------------------------
static thread_local std::string g_error;
static std::mutex g_init_mu;
static int g_device = -1;
static hipStream_t g_stream = nullptr;
static int g_num_cu = 0;
constexpr int kWavesPerCu = 16;   // resident wavefronts of sw_forward per CU

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
struct Tasks {
    std::vector<int32_t> index;   // pair numbers, grouped per task
    std::vector<int32_t> begin;   // task boundaries into index
    std::vector<int32_t> query;   // query of each task
    std::vector<int64_t> cost;    // rows x widest subject
};

// Group pairs by query, sort each query's pairs by subject length (longest
// first) and cut them into tasks of at most 128 pairs: the two halves of a
// lane and the lanes of a wavefront then walk subjects of similar length.
static void build_tasks(const std::vector<int32_t>& pairs,
                        const std::vector<int32_t>& pair_query,
                        const std::vector<int32_t>& pair_subject,
                        const std::vector<int32_t>& q_len,
                        const std::vector<int32_t>& s_len, Tasks& t) {
    std::vector<int32_t> order(pairs);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        if (pair_query[a] != pair_query[b]) return pair_query[a] < pair_query[b];
        return s_len[pair_subject[a]] > s_len[pair_subject[b]];
    });
    t.index.clear(); t.begin.clear(); t.query.clear(); t.cost.clear();
    size_t i = 0;
    while (i < order.size()) {
        const int32_t q = pair_query[order[i]];
        size_t j = i;
        while (j < order.size() && pair_query[order[j]] == q && j - i < 128) ++j;
        t.begin.push_back((int32_t)t.index.size());
        t.query.push_back(q);
        t.cost.push_back((int64_t)q_len[q] * s_len[pair_subject[order[i]]]);
        for (size_t k = i; k < j; ++k) t.index.push_back(order[k]);
        i = j;
    }
    t.begin.push_back((int32_t)t.index.size());
}

}  // namespace tpz

using namespace tpz;

struct tpz_sw_batch {
    int32_t n_queries = 0, n_subjects = 0, n_pairs = 0;
    int32_t gap_open = 11, gap_extend = 1, bias = 0;
    int32_t max_subject = 0;
    std::vector<int32_t> q_len, s_len, pair_query, pair_subject;
    DeviceBuffer d_q_res, d_q_off, d_q_len, d_s_res, d_s_off, d_s_len;
    DeviceBuffer d_pair_query, d_pair_subject, d_matrix;
    DeviceBuffer d_score, d_ref, d_read;            // results
    DeviceBuffer d_workspace, d_counter;
    // tasks grouped per query (plain schedule) and by cost (persistent)
    Tasks tasks_plain, tasks_cost;
    DeviceBuffer d_plain_index, d_plain_begin, d_cost_index, d_cost_begin;
    size_t ws_slots = 0;
    int64_t ws_words_per_slot = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms = 0;
};

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
    if (!(gap_extend >= 0 && gap_open > gap_extend && gap_open <= 255))
        return fail("gap penalties must satisfy 0 <= gap_extend < gap_open <= 255 "
                    "(got %d/%d)", gap_open, gap_extend);
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
    for (int32_t k = 0; k < n_queries; ++k)
        if (q_len[k] < 0) { delete b; return fail("query %d has negative length", k); }
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
    // subjects: each starts on a 4-byte boundary
    std::vector<int64_t> so(n_subjects);
    int64_t stot = 0;
    for (int32_t k = 0; k < n_subjects; ++k) {
        if (s_len[k] < 0) { delete b; return fail("subject %d has negative length", k); }
        so[k] = stot;
        stot += (s_len[k] + 3) & ~3;
        b->max_subject = std::max(b->max_subject, s_len[k]);
    }
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
    rc |= b->d_score.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_ref.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_read.alloc((size_t)n_pairs * 4 + 4);
    rc |= b->d_counter.alloc(64);
    if (rc) { delete b; return -1; }

    // tasks: per query (plain) and the same tasks most expensive first
    std::vector<int32_t> all(n_pairs);
    std::iota(all.begin(), all.end(), 0);
    build_tasks(all, b->pair_query, b->pair_subject, b->q_len, b->s_len, b->tasks_plain);
    {
        const Tasks& tp = b->tasks_plain;
        std::vector<int32_t> perm(tp.cost.size());
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int32_t x, int32_t y) { return tp.cost[x] > tp.cost[y]; });
        Tasks& tc = b->tasks_cost;
        for (int32_t p : perm) {
            tc.begin.push_back((int32_t)tc.index.size());
            tc.query.push_back(tp.query[p]);
            tc.cost.push_back(tp.cost[p]);
            for (int32_t k = tp.begin[p]; k < tp.begin[p + 1]; ++k) tc.index.push_back(tp.index[k]);
        }
        tc.begin.push_back((int32_t)tc.index.size());
    }
    rc |= b->d_plain_index.upload(b->tasks_plain.index.data(), b->tasks_plain.index.size() * 4);
    rc |= b->d_plain_begin.upload(b->tasks_plain.begin.data(), b->tasks_plain.begin.size() * 4);
    rc |= b->d_cost_index.upload(b->tasks_cost.index.data(), b->tasks_cost.index.size() * 4);
    rc |= b->d_cost_begin.upload(b->tasks_cost.begin.data(), b->tasks_cost.begin.size() * 4);

    // workspace: one slot (bottom H and F of a chunk, 2 words per column
    // and lane) per concurrently running wavefront
    b->ws_words_per_slot = (int64_t)2 * (b->max_subject + 2) * 64;
    size_t max_tasks_per_query = 0;
    {
        std::vector<int32_t> per_q(n_queries, 0);
        for (size_t t = 0; t + 1 < b->tasks_plain.begin.size(); ++t)
            max_tasks_per_query = std::max<size_t>(max_tasks_per_query, ++per_q[b->tasks_plain.query[t]]);
    }
    const size_t n_tasks_all = b->tasks_plain.begin.size() - 1;
    const size_t persistent_slots =
        std::min<size_t>((size_t)std::max(num_cu(), 1) * kWavesPerCu, n_tasks_all);
    b->ws_slots = std::max<size_t>(std::max(persistent_slots, max_tasks_per_query), 1);
    rc |= b->d_workspace.alloc(b->ws_slots * (size_t)b->ws_words_per_slot * 4);
    for (auto& e : b->ev) rc |= check(hipEventCreate(&e), "hipEventCreate");
    rc |= check(hipStreamSynchronize(stream()), "upload sync");
    if (rc) { delete b; return -1; }
    *out = b;
    return 0;
}

static SwKernelArgs make_args(tpz_sw_batch* b) {
    SwKernelArgs a{};
    a.q_res = (const int8_t*)b->d_q_res.ptr;
    a.q_off = (const int64_t*)b->d_q_off.ptr;
    a.q_len = (const int32_t*)b->d_q_len.ptr;
    a.s_res = (const int8_t*)b->d_s_res.ptr;
    a.s_off = (const int64_t*)b->d_s_off.ptr;
    a.s_len = (const int32_t*)b->d_s_len.ptr;
    a.pair_query = (const int32_t*)b->d_pair_query.ptr;
    a.pair_subject = (const int32_t*)b->d_pair_subject.ptr;
    a.matrix = (const int8_t*)b->d_matrix.ptr;
    a.gap_open = b->gap_open;
    a.gap_extend = b->gap_extend;
    a.workspace = (uint32_t*)b->d_workspace.ptr;
    a.ws_words_per_slot = b->ws_words_per_slot;
    a.out_score = (int32_t*)b->d_score.ptr;
    a.out_ref = (int32_t*)b->d_ref.ptr;
    a.out_read = (int32_t*)b->d_read.ptr;
    return a;
}

int tpz_sw_batch_run(tpz_sw_batch* b, int32_t schedule) {
    if (!b) return fail("null batch");
    if (schedule != TPZ_SCHED_PLAIN && schedule != TPZ_SCHED_PERSISTENT)
        return fail("unknown schedule %d", schedule);
    SwKernelArgs a = make_args(b);
    if (check(hipEventRecord(b->ev[0], stream()), "event")) return -1;
    if (schedule == TPZ_SCHED_PERSISTENT) {
        const Tasks& t = b->tasks_cost;
        const int n_tasks = (int)t.begin.size() - 1;
        if (n_tasks > 0) {
            a.pair_index = (const int32_t*)b->d_cost_index.ptr;
            a.task_begin = (const int32_t*)b->d_cost_begin.ptr;
            a.n_tasks = n_tasks;
            a.task_counter = (int32_t*)b->d_counter.ptr;
            if (check(hipMemsetAsync(b->d_counter.ptr, 0, 4, stream()), "memset counter")) return -1;
            const int grid = (int)std::min<size_t>((size_t)n_tasks, b->ws_slots);
            if (check(tpz_launch_sw_forward(&a, grid, stream()), "sw_forward (persistent)")) return -1;
        }
    } else {
        // one launch per query, one workgroup per task of that query
        const Tasks& t = b->tasks_plain;
        const int n_tasks = (int)t.begin.size() - 1;
        a.pair_index = (const int32_t*)b->d_plain_index.ptr;
        a.task_counter = nullptr;
        int k = 0;
        while (k < n_tasks) {
            int e = k;
            while (e < n_tasks && t.query[e] == t.query[k]) ++e;
            a.task_begin = (const int32_t*)b->d_plain_begin.ptr + k;
            a.n_tasks = e - k;
            if (check(tpz_launch_sw_forward(&a, e - k, stream()), "sw_forward (plain)")) return -1;
            k = e;
        }
    }
    if (check(hipEventRecord(b->ev[1], stream()), "event")) return -1;
    if (check(hipStreamSynchronize(stream()), "sw_forward")) return -1;
    float ms = 0;
    if (check(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]), "elapsed")) return -1;
    b->ms = ms;
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

}  // extern "C"
