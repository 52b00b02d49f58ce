// Served form: a request loop over resident tables with continuous batching.
//
// Requests (table, op, n keys with their own device buffers) arrive open-loop.  A scheduler thread
// keeps one FIFO per table and, whenever one of its stream slots is free, forms ONE persistent
// launch (htab_batch_run_segments) from the longest same-operation prefix of every idle table's
// FIFO, oldest head first, up to an element budget per launch.  A table is in at most one launch at
// a time, so per table the requests take effect in arrival order (finds see exactly the inserts and
// erases that arrived before them) while launches on different tables overlap on the device.
// Admission: an insert is taken only if the table's known size plus in-flight inserts plus its n
// fits the table's limit (otherwise it is refused with HTAB_E_LIMIT and has no effect); the element
// budget bounds the work in flight.  After each launch a small kernel reads the sizes of the tables
// it touched, which the scheduler uses for admission once the launch's event completes.
//
// No reference counterpart: the reference serves one blocking bulk call at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "htab_internal.hpp"

using namespace htab_host;

namespace
{

template <int Unused = 0>
__global__ void
k_table_sizes(const int64_t* const* stripes, int32_t ntables, uint32_t nstripes_max, int64_t* out)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= ntables) return;
    int64_t s = 0;
    for (uint32_t i = 0; i < nstripes_max; ++i) s += stripes[t][static_cast<uint64_t>(i) * htab_dev::kStripeStride];
    out[t] = s;
}

int64_t
now_ns(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
}

struct slot
{
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    htab_batch* batch = nullptr;
    htab_segment* segs = nullptr; // pinned
    int64_t* sizes = nullptr;     // pinned, ntables
    int64_t* sizes_dev = nullptr;
    std::vector<int64_t> reqs;     // request indices in flight
    std::vector<int32_t> tables;   // tables in flight
    int64_t elems = 0;
    bool busy = false;
};

} // namespace

struct htab_server
{
    std::vector<htab*> tables;
    std::vector<int64_t> limit;
    std::vector<slot> slots;
    const int64_t** stripes_dev = nullptr; // device array of each table's stripes pointer
    int64_t max_launch_elems = 0;
    int32_t max_segs = 0;
};

extern "C" {

int
htab_server_create(htab* const* tables, int32_t ntables, int32_t nstreams, int64_t max_inflight_elems,
                   htab_server** out)
{
    if (!out || !tables || ntables <= 0 || nstreams <= 0 || nstreams > 16 || max_inflight_elems <= 0)
        return fail(HTAB_E_ARG, "htab_server_create: bad argument");
    *out = nullptr;
    htab_server* s = new (std::nothrow) htab_server();
    if (!s) return fail(HTAB_E_OOM, "htab_server_create: host allocation");
    s->tables.assign(tables, tables + ntables);
    for (int32_t t = 0; t < ntables; ++t) s->limit.push_back(tables[t]->info.limit);
    s->max_launch_elems = std::max<int64_t>(1, max_inflight_elems / nstreams);
    s->max_segs = 4096;
    s->slots.resize(static_cast<size_t>(nstreams));
    auto cleanup = [&](int rc) {
        htab_server_destroy(s);
        return rc;
    };
    std::vector<const int64_t*> sp(static_cast<size_t>(ntables));
    for (int32_t t = 0; t < ntables; ++t) sp[static_cast<size_t>(t)] = tables[t]->view.stripes;
    if (hipMalloc(&s->stripes_dev, sizeof(int64_t*) * static_cast<size_t>(ntables)) != hipSuccess)
        return cleanup(fail(HTAB_E_OOM, "htab_server_create: device allocation"));
    if (hipMemcpy(s->stripes_dev, sp.data(), sizeof(int64_t*) * sp.size(), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(HTAB_E_HIP, "htab_server_create: copy"));
    for (auto& sl : s->slots)
    {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&sl.segs), sizeof(htab_segment) * static_cast<size_t>(s->max_segs),
                          hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&sl.sizes), sizeof(int64_t) * static_cast<size_t>(ntables),
                          hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&sl.sizes_dev, sizeof(int64_t) * static_cast<size_t>(ntables)) != hipSuccess)
            return cleanup(fail(HTAB_E_OOM, "htab_server_create: stream slot"));
        const int rc = htab_batch_create(tables, ntables, &sl.batch);
        if (rc != HTAB_OK) return cleanup(rc);
    }
    *out = s;
    return HTAB_OK;
}

int
htab_server_destroy(htab_server* s)
{
    if (!s) return fail(HTAB_E_ARG, "htab_server_destroy: null handle");
    for (auto& sl : s->slots)
    {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.batch) htab_batch_destroy(sl.batch);
        if (sl.segs) (void)hipHostFree(sl.segs);
        if (sl.sizes) (void)hipHostFree(sl.sizes);
        if (sl.sizes_dev) (void)hipFree(sl.sizes_dev);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->stripes_dev) (void)hipFree(s->stripes_dev);
    delete s;
    return HTAB_OK;
}

int
htab_server_run(htab_server* s, const htab_segment* reqs, const int64_t* arrival_ns, int64_t nreq,
                int64_t* submit_ns, int64_t* done_ns, int32_t* status, htab_server_stats* stats)
{
    if (!s || (nreq > 0 && (!reqs || !arrival_ns || !submit_ns || !done_ns || !status)) || nreq < 0)
        return fail(HTAB_E_ARG, "htab_server_run: bad argument");
    const int32_t ntables = static_cast<int32_t>(s->tables.size());
    for (int64_t i = 0; i < nreq; ++i)
    {
        if (reqs[i].table < 0 || reqs[i].table >= ntables || reqs[i].op < HTAB_OP_INSERT || reqs[i].op > HTAB_OP_ERASE ||
            reqs[i].n < 0 || (reqs[i].n > 0 && !reqs[i].keys))
            return fail(HTAB_E_ARG, "htab_server_run: bad request");
        status[i] = 1; // pending
    }
    // table sizes at start (sequentialises with any earlier work on the tables)
    std::vector<int64_t> known(static_cast<size_t>(ntables), 0), inflight_ins(static_cast<size_t>(ntables), 0);
    for (int32_t t = 0; t < ntables; ++t)
    {
        const int rc = htab_size(s->tables[static_cast<size_t>(t)], &known[static_cast<size_t>(t)], nullptr);
        if (rc != HTAB_OK) return rc;
    }
    std::vector<std::deque<int64_t>> fifo(static_cast<size_t>(ntables));
    std::vector<char> busy(static_cast<size_t>(ntables), 0);
    std::mutex mu;
    std::atomic<int64_t> submitted{0};
    std::atomic<int> error{HTAB_OK};
    int64_t launches = 0, segments = 0, max_inflight = 0;
    const auto t0 = std::chrono::steady_clock::now();

    // submitter: open-loop arrivals at their scheduled times
    std::thread submitter([&]() {
        for (int64_t i = 0; i < nreq; ++i)
        {
            while (now_ns(t0) < arrival_ns[i]) std::this_thread::yield();
            std::lock_guard<std::mutex> lk(mu);
            submit_ns[i] = now_ns(t0);
            fifo[static_cast<size_t>(reqs[i].table)].push_back(i);
            submitted.store(i + 1, std::memory_order_release);
        }
    });

    int64_t finished = 0;
    std::vector<int32_t> order;
    while (finished < nreq && error.load() == HTAB_OK)
    {
        bool progressed = false;
        // 1. retire completed launches
        for (auto& sl : s->slots)
        {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess)
            {
                error.store(fail_hip(q, "htab_server_run: launch"));
                break;
            }
            const int64_t t = now_ns(t0);
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t r : sl.reqs)
            {
                done_ns[r] = t;
                status[r] = HTAB_OK;
                if (reqs[r].op == HTAB_OP_INSERT) inflight_ins[static_cast<size_t>(reqs[r].table)] -= reqs[r].n;
            }
            for (int32_t tb : sl.tables)
            {
                known[static_cast<size_t>(tb)] = sl.sizes[tb];
                busy[static_cast<size_t>(tb)] = 0;
            }
            finished += static_cast<int64_t>(sl.reqs.size());
            sl.reqs.clear();
            sl.tables.clear();
            sl.busy = false;
            progressed = true;
        }
        // 2. form launches on free slots
        for (auto& sl : s->slots)
        {
            if (sl.busy) continue;
            std::lock_guard<std::mutex> lk(mu);
            order.clear();
            for (int32_t t = 0; t < ntables; ++t)
                if (!busy[static_cast<size_t>(t)] && !fifo[static_cast<size_t>(t)].empty()) order.push_back(t);
            if (order.empty()) break;
            std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                return submit_ns[fifo[static_cast<size_t>(a)].front()] < submit_ns[fifo[static_cast<size_t>(b)].front()];
            });
            int32_t nseg = 0;
            int64_t elems = 0;
            for (int32_t t : order)
            {
                auto& q = fifo[static_cast<size_t>(t)];
                const int op = reqs[q.front()].op;
                bool took = false;
                while (!q.empty() && reqs[q.front()].op == op && nseg < s->max_segs)
                {
                    const int64_t r = q.front();
                    const int64_t n = reqs[r].n;
                    if (took && elems + n > s->max_launch_elems) break;
                    if (op == HTAB_OP_INSERT &&
                        known[static_cast<size_t>(t)] + inflight_ins[static_cast<size_t>(t)] + n > s->limit[static_cast<size_t>(t)])
                    {
                        // cannot be guaranteed to fit: refused, no effect on the table
                        q.pop_front();
                        status[r] = HTAB_E_LIMIT;
                        done_ns[r] = now_ns(t0);
                        ++finished;
                        continue;
                    }
                    q.pop_front();
                    sl.segs[nseg++] = reqs[r];
                    sl.reqs.push_back(r);
                    if (op == HTAB_OP_INSERT) inflight_ins[static_cast<size_t>(t)] += n;
                    elems += n;
                    took = true;
                }
                if (took)
                {
                    busy[static_cast<size_t>(t)] = 1;
                    sl.tables.push_back(t);
                }
                if (elems >= s->max_launch_elems || nseg >= s->max_segs) break;
            }
            if (nseg == 0) continue;
            int rc = htab_batch_run_segments(sl.batch, sl.segs, nseg, sl.stream);
            if (rc == HTAB_OK)
            {
                hipLaunchKernelGGL(k_table_sizes<>, dim3((ntables + 255) / 256), dim3(256), 0, sl.stream, s->stripes_dev,
                                   ntables, htab_dev::kMaxStripes, sl.sizes_dev);
                rc = launch_status("htab_server_run: sizes");
            }
            if (rc == HTAB_OK && hipMemcpyAsync(sl.sizes, sl.sizes_dev, sizeof(int64_t) * static_cast<size_t>(ntables),
                                                hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
                rc = fail(HTAB_E_HIP, "htab_server_run: size copy");
            if (rc == HTAB_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess)
                rc = fail(HTAB_E_HIP, "htab_server_run: event");
            if (rc != HTAB_OK)
            {
                error.store(rc);
                break;
            }
            sl.busy = true;
            sl.elems = elems;
            ++launches;
            segments += nseg;
            int64_t inflight = 0;
            for (auto& x : s->slots) inflight += x.busy ? 1 : 0;
            max_inflight = std::max(max_inflight, inflight);
            progressed = true;
        }
        if (!progressed) std::this_thread::yield();
    }
    submitter.join();
    for (auto& sl : s->slots) (void)hipStreamSynchronize(sl.stream);
    if (stats)
    {
        stats->launches = launches;
        stats->segments = segments;
        stats->max_concurrent_launches = max_inflight;
        stats->wall_ns = now_ns(t0);
    }
    return error.load();
}

int
htab_server_replay_plain(htab_server* s, const htab_segment* reqs, const int32_t* status, int64_t nreq, int blocking,
                         void* stream)
{
    if (!s || (nreq > 0 && !reqs) || nreq < 0) return fail(HTAB_E_ARG, "htab_server_replay_plain: bad argument");
    for (int64_t i = 0; i < nreq; ++i)
    {
        if (status && status[i] != HTAB_OK) continue; // refused by admission: never applied
        const htab_segment& r = reqs[i];
        if (r.table < 0 || r.table >= static_cast<int32_t>(s->tables.size()))
            return fail(HTAB_E_ARG, "htab_server_replay_plain: bad table");
        htab* h = s->tables[static_cast<size_t>(r.table)];
        int rc = HTAB_OK;
        if (r.op == HTAB_OP_INSERT)
            rc = htab_insert_kv(h, r.keys, r.vals, r.n, r.flags, stream);
        else if (r.op == HTAB_OP_FIND)
            rc = htab_find(h, r.keys, r.n, r.flags, r.vals_out, stream);
        else
            rc = htab_erase(h, r.keys, r.n, r.flags, stream);
        if (rc != HTAB_OK) return rc;
        if (blocking && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
            return fail(HTAB_E_HIP, "htab_server_replay_plain: sync");
    }
    return HTAB_OK;
}

} // extern "C"
