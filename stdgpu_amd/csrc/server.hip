// Served form: a request loop over resident tables with continuous batching.
//
// Requests (table, op, n keys with their own device buffers) arrive open-loop.  A scheduler thread
// keeps one FIFO per table.  While fewer than `depth` launches are in flight it forms the next
// launch from the longest same-operation run at the head of every table's FIFO (oldest head first,
// up to an element budget) and enqueues it at once: launches are not held back until the tables
// they touch are idle.  Per table the order is kept on the device instead: a launch goes onto the
// stream with the fewest launches in flight and waits (hipStreamWaitEvent) for the last launch of
// each of its tables that sits on another stream, so a table's requests take effect in arrival order
// while launches of unrelated tables run side by side.  Finds therefore see exactly the inserts and
// erases that arrived before them.
//
// A launch is one kernel per 96 segments (k_serve), whose segments travel in the kernel arguments,
// then, for tables near their limit, one single-workgroup kernel that stores their sizes into
// host-mapped memory, and an event record: no copies and no planning kernel.
//
// Admission: inserts are coalesced while the table's known size plus the inserts in flight plus
// their n stay within the table's limit.  The known size is an upper bound: the last size read back
// plus every key inserted since (erases ignored, keys already present counted).  A launch reads back
// the exact size only of a table whose bound has passed 7/8 of its limit, so tables far from their
// limit cost no size kernel (and no kernel boundary) per launch.  An insert that may reach the limit
// waits until the table has no launch in flight and then runs with no other request of that table
// (its size is read back); its kernel ranks every new key against the exact limit, so it gets what a
// plain call would (keys past the limit are not inserted) and no request is ever refused.
//
// No reference counterpart: the reference serves one blocking bulk call at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "htab_internal.hpp"

using namespace htab_host;

namespace
{

int64_t
now_ns(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
}

struct launch_rec
{
    int stream = -1;
    hipEvent_t done = nullptr;
    std::vector<int64_t> reqs;   // request indices
    std::vector<int32_t> tables; // distinct tables
    std::vector<int64_t> ins_n;  // per table of `tables`: keys its inserts carry in this launch
    std::vector<uint8_t> sized;  // per table of `tables`: this launch reads its size back
    bool live = false;
};

constexpr int32_t kMaxLaunchSegs = 1024;
constexpr int kLaunchesPerStream = 2;

} // namespace

struct htab_server
{
    std::vector<htab*> tables;
    std::vector<int64_t> limit;
    std::vector<hipStream_t> streams;
    int64_t* sizes_host = nullptr;         // pinned, host-mapped, one entry per table
    int64_t* sizes_dev = nullptr;          // device address of sizes_host
    htab_batch* batch = nullptr;           // device descriptors of the tables
    std::vector<launch_rec> launches;      // depth records
    int64_t max_launch_elems = 0;
};

extern "C" {

int
htab_server_create(htab* const* tables, int32_t ntables, int32_t nstreams, int64_t max_inflight_elems,
                   htab_server** out)
{
    if (!out || !tables || ntables <= 0 || ntables > 32767 || nstreams <= 0 || nstreams > 16 || max_inflight_elems <= 0)
        return fail(HTAB_E_ARG, "htab_server_create: bad argument");
    *out = nullptr;
    htab_server* s = new (std::nothrow) htab_server();
    if (!s) return fail(HTAB_E_OOM, "htab_server_create: host allocation");
    auto cleanup = [&](int rc) {
        htab_server_destroy(s);
        return rc;
    };
    s->tables.assign(tables, tables + ntables);
    for (int32_t t = 0; t < ntables; ++t)
    {
        if (!tables[t]) return cleanup(fail(HTAB_E_ARG, "htab_server_create: null table"));
        s->limit.push_back(tables[t]->info.limit);
    }
    const int depth = kLaunchesPerStream * nstreams;
    s->max_launch_elems = std::max<int64_t>(1, max_inflight_elems / depth);
    int rc = htab_batch_create(tables, ntables, &s->batch);
    if (rc != HTAB_OK) return cleanup(rc);
    if (hipHostMalloc(reinterpret_cast<void**>(&s->sizes_host), sizeof(int64_t) * static_cast<size_t>(ntables),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return cleanup(fail(HTAB_E_OOM, "htab_server_create: pinned allocation"));
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&s->sizes_dev), s->sizes_host, 0) != hipSuccess)
        return cleanup(fail(HTAB_E_HIP, "htab_server_create: mapped size array"));
    s->streams.assign(static_cast<size_t>(nstreams), nullptr);
    for (auto& st : s->streams)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(HTAB_E_HIP, "htab_server_create: stream"));
    s->launches.resize(static_cast<size_t>(depth));
    for (auto& l : s->launches)
        if (hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(HTAB_E_HIP, "htab_server_create: event"));
    *out = s;
    return HTAB_OK;
}

int
htab_server_destroy(htab_server* s)
{
    if (!s) return fail(HTAB_E_ARG, "htab_server_destroy: null handle");
    for (auto st : s->streams)
        if (st) (void)hipStreamSynchronize(st);
    for (auto& l : s->launches)
        if (l.done) (void)hipEventDestroy(l.done);
    for (auto st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    if (s->batch) htab_batch_destroy(s->batch);
    if (s->sizes_host) (void)hipHostFree(s->sizes_host);
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
            reqs[i].n < 0 || reqs[i].n > 0xFFFFFFFFLL || (reqs[i].n > 0 && !reqs[i].keys) ||
            !segment_aligned(s->batch->kind, reqs[i]))
            return fail(HTAB_E_ARG, "htab_server_run: bad request");
        status[i] = 1; // pending
    }
    // table sizes at start (sequentialises with any earlier work on the tables)
    std::vector<int64_t> known(static_cast<size_t>(ntables), 0), inflight_ins(static_cast<size_t>(ntables), 0);
    for (int32_t t = 0; t < ntables; ++t)
    {
        const int rc = htab_size(s->tables[static_cast<size_t>(t)], &known[static_cast<size_t>(t)], nullptr);
        if (rc != HTAB_OK) return rc;
        s->sizes_host[t] = known[static_cast<size_t>(t)];
    }
    std::vector<int32_t> table_launches(static_cast<size_t>(ntables), 0); // launches in flight per table
    std::vector<int32_t> last_launch(static_cast<size_t>(ntables), -1);  // latest launch per table
    std::vector<int32_t> stream_launches(s->streams.size(), 0);
    std::vector<std::deque<int64_t>> fifo(static_cast<size_t>(ntables));
    std::mutex mu;
    int error = HTAB_OK;
    int64_t launches = 0, segments = 0, max_inflight = 0, live = 0;
    const int32_t depth = static_cast<int32_t>(s->launches.size());
    const auto t0 = std::chrono::steady_clock::now();

    // submitter: open-loop arrivals at their scheduled times
    std::thread submitter([&]() {
        for (int64_t i = 0; i < nreq; ++i)
        {
            while (now_ns(t0) < arrival_ns[i]) std::this_thread::yield();
            std::lock_guard<std::mutex> lk(mu);
            submit_ns[i] = now_ns(t0);
            fifo[static_cast<size_t>(reqs[i].table)].push_back(i);
        }
    });

    int64_t finished = 0;
    std::vector<int32_t> order;
    std::vector<htab_segment> segs;
    std::vector<int32_t> waits;
    htab_dev::serve_args args{};
    htab_dev::serve_tables tabs{};
    std::vector<int32_t> sized;
    while (finished < nreq && error == HTAB_OK)
    {
        bool progressed = false;
        // 1. retire completed launches
        for (int32_t li = 0; li < depth && error == HTAB_OK; ++li)
        {
            launch_rec& l = s->launches[static_cast<size_t>(li)];
            if (!l.live) continue;
            const hipError_t q = hipEventQuery(l.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess)
            {
                error = fail_hip(q, "htab_server_run: launch");
                break;
            }
            const int64_t t = now_ns(t0);
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t r : l.reqs)
            {
                done_ns[r] = t;
                status[r] = HTAB_OK;
                if (reqs[r].op == HTAB_OP_INSERT) inflight_ins[static_cast<size_t>(reqs[r].table)] -= reqs[r].n;
            }
            for (size_t k = 0; k < l.tables.size(); ++k)
            {
                const int32_t tb = l.tables[k];
                // a launch that read the size back gives it exactly; otherwise the bound grows by the
                // launch's inserts (the table's launches complete in their order)
                if (l.sized[k])
                    known[static_cast<size_t>(tb)] = __atomic_load_n(s->sizes_host + tb, __ATOMIC_ACQUIRE);
                else
                    known[static_cast<size_t>(tb)] += l.ins_n[k];
                --table_launches[static_cast<size_t>(tb)];
                if (last_launch[static_cast<size_t>(tb)] == li) last_launch[static_cast<size_t>(tb)] = -1;
            }
            finished += static_cast<int64_t>(l.reqs.size());
            --stream_launches[static_cast<size_t>(l.stream)];
            l.reqs.clear();
            l.tables.clear();
            l.ins_n.clear();
            l.sized.clear();
            l.live = false;
            --live;
            progressed = true;
        }
        // 2. form launches while the pipeline has room
        while (live < depth && error == HTAB_OK)
        {
            int32_t li = 0;
            while (s->launches[static_cast<size_t>(li)].live) ++li;
            launch_rec& l = s->launches[static_cast<size_t>(li)];
            int32_t sid = 0;
            for (int32_t k = 1; k < static_cast<int32_t>(s->streams.size()); ++k)
                if (stream_launches[static_cast<size_t>(k)] < stream_launches[static_cast<size_t>(sid)]) sid = k;
            segs.clear();
            int64_t elems = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                order.clear();
                for (int32_t t = 0; t < ntables; ++t)
                    if (!fifo[static_cast<size_t>(t)].empty()) order.push_back(t);
                std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                    return fifo[static_cast<size_t>(a)].front() < fifo[static_cast<size_t>(b)].front();
                });
                for (int32_t t : order)
                {
                    auto& q = fifo[static_cast<size_t>(t)];
                    const int op = reqs[q.front()].op;
                    bool took = false, alone_run = false;
                    int64_t run_ins = 0;
                    while (!q.empty() && reqs[q.front()].op == op && static_cast<int32_t>(segs.size()) < kMaxLaunchSegs)
                    {
                        const int64_t r = q.front();
                        const int64_t n = reqs[r].n;
                        if (took && elems + n > s->max_launch_elems) break;
                        bool alone = false;
                        if (op == HTAB_OP_INSERT &&
                            known[static_cast<size_t>(t)] + inflight_ins[static_cast<size_t>(t)] + n >
                                    s->limit[static_cast<size_t>(t)])
                        {
                            // the request may reach the table's limit (n counts keys that may be
                            // present already): wait for the table's launches in flight, then run it
                            // with no other request of the table, so earlier requests take their
                            // slots first, as in the in-order plain calls; the insert kernel ranks
                            // each key against the exact limit
                            if (took || table_launches[static_cast<size_t>(t)] > 0) break;
                            alone = true;
                        }
                        q.pop_front();
                        segs.push_back(reqs[r]);
                        l.reqs.push_back(r);
                        if (op == HTAB_OP_INSERT)
                        {
                            inflight_ins[static_cast<size_t>(t)] += n;
                            run_ins += n;
                        }
                        elems += n;
                        took = true;
                        alone_run = alone_run || alone;
                        if (alone) break;
                    }
                    if (took)
                    {
                        l.tables.push_back(t);
                        l.ins_n.push_back(run_ins);
                        const __int128 bound = static_cast<__int128>(known[static_cast<size_t>(t)]) +
                                               inflight_ins[static_cast<size_t>(t)];
                        l.sized.push_back(op != HTAB_OP_FIND &&
                                                  (alone_run || bound * 8 > static_cast<__int128>(s->limit[static_cast<size_t>(t)]) * 7)
                                          ? 1
                                          : 0);
                    }
                    if (elems >= s->max_launch_elems || static_cast<int32_t>(segs.size()) >= kMaxLaunchSegs) break;
                }
            }
            if (segs.empty()) break;
            hipStream_t st = s->streams[static_cast<size_t>(sid)];
            // per-table order: wait for the latest launch of each table that sits on another stream
            waits.clear();
            for (int32_t t : l.tables)
            {
                const int32_t p = last_launch[static_cast<size_t>(t)];
                if (p >= 0 && s->launches[static_cast<size_t>(p)].stream != sid &&
                    std::find(waits.begin(), waits.end(), p) == waits.end())
                    waits.push_back(p);
            }
            int rc = HTAB_OK;
            for (int32_t p : waits)
                if (hipStreamWaitEvent(st, s->launches[static_cast<size_t>(p)].done, 0) != hipSuccess)
                    rc = fail(HTAB_E_HIP, "htab_server_run: stream wait");
            for (size_t g = 0; g < segs.size() && rc == HTAB_OK; g += htab_dev::kServeSegs)
            {
                const size_t m = std::min(segs.size() - g, static_cast<size_t>(htab_dev::kServeSegs));
                uint32_t chunks = 0;
                for (size_t j = 0; j < m; ++j)
                {
                    const htab_segment& r = segs[g + j];
                    htab_dev::serve_seg& d = args.segs[j];
                    d.keys = r.keys;
                    d.vals = r.op == HTAB_OP_FIND ? r.vals_out : const_cast<void*>(r.vals);
                    d.flags = r.flags;
                    d.n = static_cast<uint32_t>(r.n);
                    d.table = static_cast<int16_t>(r.table);
                    d.op = static_cast<int8_t>(r.op);
                    d.pad = 0;
                    args.chunk_start[j] = chunks;
                    chunks += static_cast<uint32_t>((r.n + htab_dev::kServeBlock - 1) / htab_dev::kServeBlock);
                }
                args.chunk_start[m] = chunks;
                args.nsegs = static_cast<int32_t>(m);
                if (chunks == 0) continue;
                htab_call c;
                c.op = call_op::serve;
                c.b = s->batch;
                c.serve = &args;
                c.s = st;
                rc = dispatch(s->batch->kind, c);
            }
            // sizes of the tables whose run inserts or erases (a find leaves the size alone)
            sized.clear();
            for (size_t k = 0; k < l.tables.size(); ++k)
                if (l.sized[k]) sized.push_back(l.tables[k]);
            for (size_t g = 0; g < sized.size() && rc == HTAB_OK; g += htab_dev::kServeTables)
            {
                const size_t m = std::min(sized.size() - g, static_cast<size_t>(htab_dev::kServeTables));
                tabs.n = static_cast<int32_t>(m);
                for (size_t j = 0; j < m; ++j) tabs.table[j] = static_cast<int16_t>(sized[g + j]);
                htab_call c;
                c.op = call_op::serve_sizes;
                c.b = s->batch;
                c.serve_tabs = &tabs;
                c.sizes_out = s->sizes_dev;
                c.s = st;
                rc = dispatch(s->batch->kind, c);
            }
            if (rc == HTAB_OK && hipEventRecord(l.done, st) != hipSuccess)
                rc = fail(HTAB_E_HIP, "htab_server_run: event");
            if (rc != HTAB_OK)
            {
                error = rc;
                break;
            }
            for (int32_t t : l.tables)
            {
                ++table_launches[static_cast<size_t>(t)];
                last_launch[static_cast<size_t>(t)] = li;
            }
            l.stream = sid;
            l.live = true;
            ++stream_launches[static_cast<size_t>(sid)];
            ++live;
            ++launches;
            segments += static_cast<int64_t>(segs.size());
            max_inflight = std::max(max_inflight, live);
            progressed = true;
        }
        if (!progressed) std::this_thread::yield();
    }
    submitter.join();
    for (auto st : s->streams) (void)hipStreamSynchronize(st);
    for (auto& l : s->launches)
    {
        l.reqs.clear();
        l.tables.clear();
        l.live = false;
    }
    if (stats)
    {
        stats->launches = launches;
        stats->segments = segments;
        stats->max_concurrent_launches = max_inflight;
        stats->wall_ns = now_ns(t0);
    }
    return error;
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
