// Serving engine: resident StaticMap / StaticMultimap tables in HBM and a
// request loop with continuous batching.
//
//  - submit() appends a request (table, op, device input/output, count) to
//    an arrival-ordered queue and returns a ticket at once;
//  - the dispatcher thread drains the queue into *ticks*: every StaticMap
//    request that can join without breaking its table's arrival order goes
//    into the tick, and the tick runs as one persistent insert launch then
//    one find and one contains launch over all its requests (the batched
//    probe kernel), queued on the map stream without waiting for earlier
//    ticks -- requests that arrive while a tick runs are admitted into the
//    next one as tick buffers free up;
//  - StaticMultimap requests run on their own thread and stream, in ticks
//    of their own that overlap the map ticks on the device: per table at
//    most one insert (run as StaticMultimap::Insert) and then reads, and the
//    tick's reads of every table run as one batched pass (multimap_batch.h);
//  - a completion thread retires ticks in order (event wait), stamps each
//    request's completion time and publishes its result.
// Per table, requests take effect in arrival order: a tick takes a table's
// requests while they are one insert followed by reads, so every result
// equals the result of the same requests run one by one.  New tables are admitted
// while free HBM stays above the configured reserve.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_kernels.h"
#include "multimap_batch.h"
#include "mxco_common.h"
#include "serve_resident.h"

using namespace mxco;

namespace {

double now_us()
{
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    mxco_request r;
    BatchTable bt{};  // a StaticMap request's table, taken at submit
    uint64_t ticket;
    double t_submit = 0, t_done = 0;
    uint64_t result = 0;
    mxco_status status = MXCO_OK;
    bool done = false;
};

struct Table {
    int kind;  // 0 map, 1 multimap
    mxco_oa* map = nullptr;
    BatchTable bt{};  // a map's slot array, capacity and bucket count
    mxco_mm* mm = nullptr;
    uint64_t bytes = 0;
    bool live = false;
};

constexpr int kTickSlots = kServeSlots;
constexpr int kMmWorkers = 1;  // StaticMultimap tick threads (and streams)
constexpr uint32_t kMaxTickRequests = kServeMaxItems;
constexpr uint64_t kMaxMmTickKeys = 1ull << 22;  // multimap queries per tick (one larger request runs alone)

struct Tick {
    // One upload per tick: [items | chunks | zeroed insert counters].
    uint8_t* blob_host = nullptr;  // pinned
    uint8_t* blob_dev = nullptr;
    size_t blob_cap = 0;
    unsigned long long* counters_host = nullptr;  // pinned read-back of the insert counters
    uint32_t counters_cap = 0;
    hipEvent_t done = nullptr;
    // Outgrown buffers, freed by destroy: hipFree waits for the whole device
    // and would stall every tick and multimap request in flight.
    std::vector<void*> retired_host, retired_dev;
    std::vector<Req*> reqs;
    std::vector<uint32_t> insert_items;  // request index of each insert item
    uint64_t seq = 0;                    // resident mode: the tick's number
    bool busy = false;
};

}  // namespace

struct mxco_server {
    mxco_server_config cfg;
    mxco_oa_desc map_desc;
    int device = 0;
    std::vector<Table> tables;
    BatchTable* tables_dev = nullptr;
    uint32_t tables_cap = 0;
    std::mutex mu;
    std::condition_variable cv_pending, cv_done, cv_ticks, cv_mm;
    std::deque<Req*> pending;
    std::deque<Req*> mm_pending[kMmWorkers];
    std::unordered_map<uint64_t, Req*> reqs;
    uint64_t next_ticket = 1;
    Tick ticks[kTickSlots];
    std::deque<Tick*> inflight;
    hipStream_t map_stream = nullptr, mm_streams[kMmWorkers] = {};
    MmBatchWs* mm_ws[kMmWorkers] = {};
    std::thread dispatcher, completer, mm_workers[kMmWorkers];
    bool stop = false;
    std::string worker_error;
    // Resident mode: ticks go to the resident kernel through the mailbox
    // (serve_resident.h) instead of one launch pair each.  The completion
    // thread (re)launches the kernel when a posted tick is unfinished and no
    // launch is running; res_* below are guarded by `mu`.
    bool resident = false;
    ServeMailbox* mb = nullptr;  // host-coherent
    ServeCtl* ctl = nullptr;     // device
    hipStream_t res_stream = nullptr;  // CU-masked over every CU: a hardware queue of its own
    hipEvent_t res_done = nullptr;
    bool res_running = false;
    uint64_t res_posted = 0;  // seq of the last posted tick
    uint64_t res_launches = 0;
    unsigned res_grid = 0;
    uint64_t res_idle = 0, res_life = 0, res_hard = 0;  // in wall-clock ticks
    double res_ticks_per_us = 100.0;
    uint32_t res_group = 1;  // lanes per key of the resident kernel
    // statistics
    uint64_t n_ticks = 0, n_tick_requests = 0;
};

namespace {

// A stream with a hardware queue of its own, for the resident kernel: HIP
// gives a stream with a CU mask (here every CU) its own queue, while
// ordinary streams share the process's few queues, so a kernel queued on a
// stream that shares the resident kernel's queue would wait until it exits.
// Such a stream is a blocking one: work on the legacy default stream is
// ordered with it.
hipError_t queue_stream(hipStream_t* st)
{
    std::vector<uint32_t> mask(((unsigned)cu_count() + 31) / 32, 0xffffffffu);
    return hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data());
}

mxco_status grow_tick(Tick& t, size_t bytes, uint32_t counters, bool resident)
{
    // Sized at create for a full tick; a tick grows only for one request
    // larger than max_tick_keys.  The slot is idle here (its previous tick
    // retired), but the old buffers are only freed at destroy.  The resident
    // kernel reads the request list and writes the counters in place, so in
    // that mode both are host-coherent and there is no device copy.
    const unsigned host_flags = resident ? hipHostMallocCoherent : 0;
    if (bytes > t.blob_cap) {
        const size_t cap = std::max(bytes, 2 * t.blob_cap);
        uint8_t *h = nullptr, *d = nullptr;
        MXCO_HIP(hipHostMalloc((void**)&h, cap, host_flags), "server: tick buffers");
        if (!resident) {
            if (hipError_t e = hipMalloc((void**)&d, cap); e != hipSuccess) {
                (void)hipHostFree(h);
                return hip_fail(e, "server: tick buffers");
            }
        }
        if (t.blob_host) t.retired_host.push_back(t.blob_host);
        if (t.blob_dev) t.retired_dev.push_back(t.blob_dev);
        t.blob_host = h;
        t.blob_dev = d;
        t.blob_cap = cap;
    }
    if (counters > t.counters_cap) {
        const uint32_t cap = std::max(counters, 2 * t.counters_cap);
        unsigned long long* h = nullptr;
        MXCO_HIP(hipHostMalloc((void**)&h, cap * sizeof(unsigned long long), host_flags), "server: counters");
        if (t.counters_host) t.retired_host.push_back(t.counters_host);
        t.counters_host = h;
        t.counters_cap = cap;
    }
    return MXCO_OK;
}

OaArgs map_args(const mxco_server* s)
{
    OaArgs a;
    memset(&a, 0, sizeof(a));
    a.bucket = s->map_desc.bucket_size;
    a.probing = s->map_desc.probing;
    a.hash = s->map_desc.hash;
    a.key_float = s->map_desc.key_kind == MXCO_KEY_FLOAT;
    memcpy(&a.empty_key, s->cfg.empty_key, s->map_desc.key_size);
    memcpy(&a.empty_val, s->cfg.empty_value, s->map_desc.value_size);
    a.erased_key = a.empty_key;
    return a;
}

bool is_map_op(uint32_t op) { return op == MXCO_REQ_INSERT || op == MXCO_REQ_FIND || op == MXCO_REQ_CONTAINS; }

// Keys per chunk of one launch of a tick: about 16 chunks per CU, so that a
// launch of a few hundred thousand keys still spreads over every CU with a
// few keys per lane group (each a chain of dependent line reads); at least
// 64, at most the batched path's chunk.
uint32_t tick_chunk(uint64_t keys)
{
    const uint64_t target = keys / (16ull * (uint64_t)cu_count());
    uint32_t c = 64;
    while (c < kChunk && c * 2 <= target) c *= 2;
    return c;
}

struct TickLayout {
    size_t off_chunks, off_counters, bytes;
};

// One tick's upload: [items | chunks | zeroed insert counters], 16-byte aligned parts.
TickLayout tick_layout(uint64_t nitems, uint64_t nchunks, uint64_t ninsert)
{
    TickLayout l;
    l.off_chunks = (nitems * sizeof(ServeItem) + 15) & ~size_t(15);
    l.off_counters = (l.off_chunks + nchunks * sizeof(BatchChunk) + 15) & ~size_t(15);
    l.bytes = l.off_counters + 2ull * ninsert * sizeof(unsigned long long);
    return l;
}

// The most chunks a tick of at most `keys` keys in `nitems` requests cuts:
// each of its two launches cuts k keys into chunks of tick_chunk(k), which
// exceeds k / (32 * CUs) unless it is capped at 64 (then k / 64 < 32 * CUs
// too) or at kChunk; each request adds at most one partial chunk.
uint64_t tick_chunk_bound(uint64_t keys, uint64_t nitems)
{
    return 2 * std::max<uint64_t>(32ull * (uint64_t)cu_count(), keys / kChunk + 1) + nitems;
}

// Runs one tick's requests (already chosen) on the map stream: one upload,
// one insert launch, one read launch (Find and Contains), one read-back.
mxco_status launch_tick(mxco_server* s, Tick& t)
{
    std::vector<uint32_t> ins, rd;
    uint64_t ins_keys = 0, rd_keys = 0;
    for (uint32_t i = 0; i < t.reqs.size(); ++i) {
        const bool insert = t.reqs[i]->r.op == MXCO_REQ_INSERT;
        (insert ? ins : rd).push_back(i);
        (insert ? ins_keys : rd_keys) += t.reqs[i]->r.n;
    }
    // each launch cut by its own key count
    const uint32_t csz_ins = tick_chunk(ins_keys), csz_rd = tick_chunk(rd_keys);
    const uint32_t nitems = (uint32_t)t.reqs.size();
    uint32_t nchunks = 0;
    for (Req* r : t.reqs) {
        const uint32_t csz = r->r.op == MXCO_REQ_INSERT ? csz_ins : csz_rd;
        nchunks += (uint32_t)((r->r.n + csz - 1) / csz);
    }
    const uint32_t ninsert = (uint32_t)ins.size();
    const TickLayout lay = tick_layout(nitems, nchunks, ninsert);
    const size_t off_chunks = lay.off_chunks, off_counters = lay.off_counters, bytes = lay.bytes;
    if (mxco_status e = grow_tick(t, bytes, 2 * ninsert + 2, false)) return e;
    ServeItem* items = (ServeItem*)t.blob_host;
    BatchChunk* chunks = (BatchChunk*)(t.blob_host + off_chunks);
    memset(t.blob_host + off_counters, 0, 2ull * ninsert * sizeof(unsigned long long));
    uint32_t item = 0, chunk = 0, first_read = 0;
    t.insert_items.clear();
    for (int k = 0; k < 2; ++k) {
        if (k == 1) first_read = chunk;
        const uint32_t csz = k == 0 ? csz_ins : csz_rd;
        for (uint32_t ri : (k == 0 ? ins : rd)) {
            const mxco_request& r = t.reqs[ri]->r;
            items[item] = {r.table, r.op, (const uint8_t*)r.in, (uint8_t*)r.out, r.n};
            if (k == 0) t.insert_items.push_back(ri);
            for (uint64_t st = 0; st < r.n; st += csz)
                chunks[chunk++] = {item, (uint32_t)std::min<uint64_t>(csz, r.n - st), st};
            ++item;
        }
    }
    hipStream_t st = s->map_stream;
    MXCO_HIP(hipMemcpyAsync(t.blob_dev, t.blob_host, bytes, hipMemcpyHostToDevice, st), "server: upload");
    const OaArgs a = map_args(s);
    const ServeItem* items_dev = (const ServeItem*)t.blob_dev;
    const BatchChunk* chunks_dev = (const BatchChunk*)(t.blob_dev + off_chunks);
    unsigned long long* counters_dev = (unsigned long long*)(t.blob_dev + off_counters);
    const uint32_t n_ins_chunks = first_read, n_rd_chunks = chunk - first_read;
    const unsigned cap = (unsigned)cu_count() * s->cfg.blocks_per_cu;
    if (n_ins_chunks)
        MXCO_HIP(launch_batch_any(s->map_desc.key_size, s->map_desc.value_size, s->cfg.group_size, OP_INSERT, true,
                                  a, s->tables_dev, items_dev, chunks_dev, n_ins_chunks, counters_dev,
                                  std::min(n_ins_chunks, cap), st),
                 "server: insert launch");
    if (n_rd_chunks)
        MXCO_HIP(launch_batch_any(s->map_desc.key_size, s->map_desc.value_size, s->cfg.group_size, OP_FIND, true, a,
                                  s->tables_dev, items_dev, chunks_dev + first_read, n_rd_chunks, nullptr,
                                  std::min(n_rd_chunks, cap), st),
                 "server: read launch");
    if (ninsert)
        MXCO_HIP(hipMemcpyAsync(t.counters_host, counters_dev, 2ull * ninsert * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st),
                 "server: read counters");
    MXCO_HIP(hipEventRecord(t.done, st), "server: record");
    return MXCO_OK;
}

// Keys per chunk of a resident tick phase of `keys` keys: about two chunks
// per worker, at least one key per lane group of a worker, at most kChunk.
uint32_t resident_chunk(uint64_t keys, unsigned workers, uint32_t groups)
{
    const uint64_t target = keys / (2ull * std::max(1u, workers));
    uint32_t c = groups;
    while (c < kChunk && c * 2 <= target) c *= 2;
    return c;
}

// Resident mode: writes one tick's request list (inserts first, each with
// its keys per chunk and first chunk) into its host-coherent blob and posts
// the tick into the mailbox; the resident kernel's leader copies it in.
mxco_status post_tick(mxco_server* s, Tick& t)
{
    std::vector<uint32_t> ins, rd;
    uint64_t ins_keys = 0, rd_keys = 0;
    for (uint32_t i = 0; i < t.reqs.size(); ++i) {
        const bool insert = t.reqs[i]->r.op == MXCO_REQ_INSERT;
        (insert ? ins : rd).push_back(i);
        (insert ? ins_keys : rd_keys) += t.reqs[i]->r.n;
    }
    // reads of tables without an insert in this tick first: they need not
    // wait for the tick's inserts
    std::vector<uint32_t> ins_tables;
    for (uint32_t ri : ins) ins_tables.push_back(t.reqs[ri]->r.table);
    std::sort(ins_tables.begin(), ins_tables.end());
    auto after_insert = [&](uint32_t ri) {
        return std::binary_search(ins_tables.begin(), ins_tables.end(), t.reqs[ri]->r.table);
    };
    std::stable_partition(rd.begin(), rd.end(), [&](uint32_t ri) { return !after_insert(ri); });
    const unsigned workers = s->res_grid - 1;
    const uint32_t groups = kServeThreads / s->res_group;
    const uint32_t csz_ins = resident_chunk(ins_keys, workers, groups), csz_rd = resident_chunk(rd_keys, workers, groups);
    const uint32_t nreq = (uint32_t)t.reqs.size();
    // the kernel's leader reads whole 16-byte words
    if (mxco_status e = grow_tick(t, nreq * sizeof(ServeReq) + 16, 2 * (uint32_t)ins.size() + 2, true)) return e;
    ServeReq* reqs = (ServeReq*)t.blob_host;
    uint64_t chunk = 0;
    uint32_t n_ins_chunks = 0, item = 0;
    t.insert_items.clear();
    for (int k = 0; k < 2; ++k) {
        if (k == 1) n_ins_chunks = (uint32_t)chunk;
        const uint32_t csz = k == 0 ? csz_ins : csz_rd;
        for (uint32_t ri : (k == 0 ? ins : rd)) {
            const mxco_request& r = t.reqs[ri]->r;
            if (k == 0) t.insert_items.push_back(ri);
            const uint32_t op = r.op | (k == 1 && after_insert(ri) ? kServeAfterInsert : 0u);
            const BatchTable& bt = t.reqs[ri]->bt;
            reqs[item++] = ServeReq{(const uint8_t*)r.in, (uint8_t*)r.out, r.n, bt.slots, bt.cap, bt.nbuckets,
                                    op, csz, (uint32_t)chunk, 0u};
            chunk += (r.n + csz - 1) / csz;
            if (chunk >= kServeMaxChunks)
                return fail(MXCO_LENGTH_ERROR, "server: tick too large for the resident kernel");
        }
    }
    t.seq = ++s->res_posted;
    ServeTickDesc& d = s->mb->desc[t.seq % kServeSlots];
    d.reqs = reqs;
    d.counters_host = t.counters_host;
    d.seq = t.seq;
    d.n_reqs = nreq;
    d.n_insert_items = (uint32_t)ins.size();
    d.n_ins_chunks = n_ins_chunks;
    d.n_rd_chunks = (uint32_t)chunk - n_ins_chunks;
    __atomic_store_n(&s->mb->posted, t.seq, __ATOMIC_RELEASE);
    return MXCO_OK;
}

// Resident mode, under `mu`: starts the resident kernel at tick seq_base + 1.
mxco_status res_launch(mxco_server* s, uint64_t seq_base)
{
    MXCO_HIP(hipMemsetAsync(s->ctl, 0, kServeCtlHeader, s->res_stream), "server: resident reset");
    MXCO_HIP(launch_serve_resident(s->map_desc.key_size, s->map_desc.value_size, s->res_group, map_args(s),
                                   s->mb, s->ctl, seq_base, s->res_idle, s->res_life, s->res_hard,
                                   s->res_grid, s->res_stream),
             "server: resident launch");
    MXCO_HIP(hipEventRecord(s->res_done, s->res_stream), "server: resident record");
    s->res_running = true;
    ++s->res_launches;
    return MXCO_OK;
}

// Resident mode, under `mu`, with no tick in flight: ends the resident
// kernel (it exits at once between ticks) and waits for it.
void res_quiesce(mxco_server* s)
{
    if (!s->resident || !s->res_running) return;
    __atomic_store_n(&s->mb->stop, (uint64_t)1, __ATOMIC_RELEASE);
    (void)hipEventSynchronize(s->res_done);
    __atomic_store_n(&s->mb->stop, (uint64_t)0, __ATOMIC_RELEASE);
    s->res_running = false;
}

void finish(mxco_server* s, Req* r, uint64_t result, mxco_status st)
{
    r->result = result;
    r->status = st;
    r->t_done = now_us();
    r->done = true;
}

void dispatcher_loop(mxco_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_pending.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        if (s->stop && s->pending.empty()) break;
        // a free tick slot
        Tick* t = nullptr;
        s->cv_ticks.wait(lk, [&] {
            for (Tick& x : s->ticks)
                if (!x.busy) return true;
            return s->stop;
        });
        for (Tick& x : s->ticks)
            if (!x.busy) {
                t = &x;
                break;
            }
        if (!t) break;
        // Choose the tick's requests: per table, at most one insert followed
        // by reads (two inserts in one launch would race on a shared key,
        // where arrival order says the earlier one wins).
        std::unordered_map<uint32_t, int> phase;  // 1 inserting, 2 reading, 3 blocked
        std::deque<Req*> keep;
        uint64_t keys = 0;
        t->reqs.clear();
        for (Req* r : s->pending) {
            const uint32_t tb = r->r.table;
            if (!is_map_op(r->r.op)) {
                s->mm_pending[tb % kMmWorkers].push_back(r);
                continue;
            }
            int& ph = phase[tb];
            const bool insert = r->r.op == MXCO_REQ_INSERT;
            const bool fits =
                (keys + r->r.n <= s->cfg.max_tick_keys && t->reqs.size() < kMaxTickRequests) || t->reqs.empty();
            if (ph == 3 || !fits || (insert && ph != 0)) {
                ph = 3;
                keep.push_back(r);
                continue;
            }
            ph = insert ? 1 : 2;
            keys += r->r.n;
            t->reqs.push_back(r);
        }
        s->pending.swap(keep);
        s->cv_mm.notify_all();
        if (t->reqs.empty()) continue;
        t->busy = true;
        ++s->n_ticks;
        s->n_tick_requests += t->reqs.size();
        lk.unlock();
        mxco_status st = s->resident ? post_tick(s, *t) : launch_tick(s, *t);
        lk.lock();
        if (st != MXCO_OK) {
            for (Req* r : t->reqs) finish(s, r, 0, st);
            s->worker_error = mxco_last_error();
            t->busy = false;
            s->cv_done.notify_all();
            continue;
        }
        s->inflight.push_back(t);
        s->cv_done.notify_all();  // wakes the completer
    }
}

// Resident mode: waits for the tick's done word (polling without the lock)
// and, whenever the resident kernel is not running while the tick is
// unfinished, launches it at this tick (every earlier tick has finished:
// ticks retire in order).  Called and returns with `lk` held.
hipError_t res_wait(mxco_server* s, Tick& t, std::unique_lock<std::mutex>& lk)
{
    const uint64_t* done = &s->mb->done[t.seq % kServeSlots];
    for (;;) {
        if (__atomic_load_n(done, __ATOMIC_ACQUIRE) == t.seq) return hipSuccess;
        if (s->res_running) {
            const hipError_t q = hipEventQuery(s->res_done);
            if (q != hipSuccess && q != hipErrorNotReady) return q;
            if (q == hipSuccess) s->res_running = false;
        }
        if (!s->res_running) {
            // the last launch has ended: its ticks are finished, or it gave this one up
            if (__atomic_load_n(done, __ATOMIC_ACQUIRE) == t.seq) return hipSuccess;
            // the launch gave up this tick; a number of an older tick means that
            // tick's chunks all finished after the deadline, and it completed
            if (const uint64_t ab = __atomic_load_n(&s->mb->aborted, __ATOMIC_ACQUIRE)) {
                __atomic_store_n(&s->mb->aborted, (uint64_t)0, __ATOMIC_RELEASE);
                if (ab == t.seq) return hipErrorLaunchTimeOut;
            }
            if (res_launch(s, t.seq - 1) != MXCO_OK) return hipErrorLaunchFailure;
        }
        lk.unlock();
        for (int i = 0; i < 256 && __atomic_load_n(done, __ATOMIC_ACQUIRE) != t.seq; ++i) std::this_thread::yield();
        lk.lock();
    }
}

void completer_loop(mxco_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_done.wait(lk, [&] { return s->stop || !s->inflight.empty(); });
        if (s->inflight.empty()) {
            if (s->stop) break;
            continue;
        }
        Tick* t = s->inflight.front();
        hipError_t e = hipSuccess;
        if (s->resident) {
            e = res_wait(s, *t, lk);
        } else {
            lk.unlock();
            // poll: completion is noticed within microseconds of the event
            while ((e = hipEventQuery(t->done)) == hipErrorNotReady) std::this_thread::yield();
            lk.lock();
        }
        s->inflight.pop_front();
        const double tnow = now_us();
        if (e != hipSuccess)
            s->worker_error = e == hipErrorLaunchTimeOut ? "resident kernel: a tick exceeded its deadline"
                                                         : std::string("tick failed: ") + hipGetErrorString(e);
        for (Req* r : t->reqs) {
            r->t_done = tnow;
            r->status = e == hipSuccess ? MXCO_OK : MXCO_RUNTIME_ERROR;
            r->done = true;
        }
        for (uint32_t k = 0; k < t->insert_items.size(); ++k)
            t->reqs[t->insert_items[k]]->result = t->counters_host[2 * k];
        t->reqs.clear();
        t->busy = false;
        s->cv_ticks.notify_all();
        s->cv_done.notify_all();
    }
}

// Multimap ticks of worker w's tables, on its stream.  A tick takes, per
// table, at most one insert followed by reads (a second insert, or an insert
// after a read, waits for the next tick): the inserts run one by one, then
// all reads of the tick as one batched pass, so per table every request
// takes effect in arrival order.
void mm_loop(mxco_server* s, int w)
{
    (void)hipSetDevice(s->device);
    hipStream_t stream = s->mm_streams[w];
    std::deque<Req*>& queue = s->mm_pending[w];
    std::unique_lock<std::mutex> lk(s->mu);
    std::vector<Req*> ins, rd;
    std::vector<mxco_mm*> ins_mm, rd_mm;  // taken under the lock (tables may grow meanwhile)
    std::vector<MmRead> reads;
    for (;;) {
        s->cv_mm.wait(lk, [&] { return s->stop || !queue.empty(); });
        if (queue.empty()) {
            if (s->stop) break;
            continue;
        }
        std::unordered_map<uint32_t, int> phase;  // 1 inserting, 2 reading, 3 blocked
        std::deque<Req*> keep;
        ins.clear();
        rd.clear();
        uint64_t keys = 0;
        for (Req* r : queue) {
            int& ph = phase[r->r.table];
            const bool insert = r->r.op == MXCO_REQ_MM_INSERT;
            const bool full = ins.size() + rd.size() >= kMaxTickRequests ||
                              (!insert && !rd.empty() && keys + r->r.n > kMaxMmTickKeys);
            if (ph == 3 || (insert && ph != 0) || full) {
                ph = 3;
                keep.push_back(r);
                continue;
            }
            ph = insert ? 1 : 2;
            (insert ? ins : rd).push_back(r);
            if (!insert) keys += r->r.n;
        }
        queue.swap(keep);
        ins_mm.clear();
        rd_mm.clear();
        for (Req* r : ins) ins_mm.push_back(s->tables[r->r.table].mm);
        for (Req* r : rd) rd_mm.push_back(s->tables[r->r.table].mm);
        lk.unlock();
        std::vector<std::pair<uint64_t, mxco_status>> ins_res(ins.size());
        for (size_t k = 0; k < ins.size(); ++k) {
            const mxco_request& q = ins[k]->r;
            uint64_t result = 0;
            const mxco_status st = mxco_mm_insert(ins_mm[k], q.in, nullptr, q.n, stream, &result);
            ins_res[k] = {result, st};
        }
        reads.resize(rd.size());
        for (size_t k = 0; k < rd.size(); ++k) {
            const mxco_request& q = rd[k]->r;
            reads[k] = MmRead{rd_mm[k], q.op, q.in, q.n, q.out, q.out2, q.out_capacity, 0, MXCO_OK};
        }
        const mxco_status bst = mm_batch_reads(reads.data(), (uint32_t)reads.size(), s->mm_ws[w], stream);
        const std::string berr = bst == MXCO_OK ? std::string() : mxco_last_error();
        lk.lock();
        for (size_t k = 0; k < ins.size(); ++k) {
            if (ins_res[k].second != MXCO_OK) s->worker_error = "StaticMultimap insert failed";
            finish(s, ins[k], ins_res[k].first, ins_res[k].second);
        }
        for (size_t k = 0; k < rd.size(); ++k) {
            mxco_status st = bst != MXCO_OK ? bst : reads[k].status;
            if (bst != MXCO_OK) s->worker_error = berr;
            else if (st == MXCO_LENGTH_ERROR) s->worker_error = "StaticMultimap: output capacity is insufficient";
            else if (st != MXCO_OK) s->worker_error = "StaticMultimap: null nonempty buffer";
            finish(s, rd[k], reads[k].result, st);
        }
        s->cv_done.notify_all();
    }
}

mxco_status validate(mxco_server* s, const mxco_request* r)
{
    if (r->table >= s->tables.size() || !s->tables[r->table].live)
        return fail(MXCO_INVALID_ARGUMENT, "server: unknown table");
    const int kind = s->tables[r->table].kind;
    if (kind == 0 && !is_map_op(r->op)) return fail(MXCO_INVALID_ARGUMENT, "server: not a StaticMap operation");
    if (kind == 1 && is_map_op(r->op)) return fail(MXCO_INVALID_ARGUMENT, "server: not a StaticMultimap operation");
    if (r->n && !r->in) return fail(MXCO_INVALID_ARGUMENT, "server: null input");
    if (r->n && r->op != MXCO_REQ_INSERT && r->op != MXCO_REQ_MM_INSERT && r->op != MXCO_REQ_MM_COUNT && !r->out)
        return fail(MXCO_INVALID_ARGUMENT, "server: null output");
    return MXCO_OK;
}

mxco_status admit(mxco_server* s, uint64_t bytes)
{
    size_t free_b = 0, total_b = 0;
    MXCO_HIP(hipMemGetInfo(&free_b, &total_b), "server: hipMemGetInfo");
    uint64_t used = 0;
    for (const Table& t : s->tables)
        if (t.live) used += t.bytes;
    if (s->cfg.hbm_budget_bytes && used + bytes > s->cfg.hbm_budget_bytes)
        return fail(MXCO_OUT_OF_MEMORY, "server: admission refused (table budget)");
    if ((uint64_t)free_b < bytes + s->cfg.hbm_reserve_bytes)
        return fail(MXCO_OUT_OF_MEMORY, "server: admission refused (free HBM)");
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_server_create(const mxco_server_config* cfg, const mxco_oa_desc* map_desc,
                                        mxco_server** out)
{
    if (!cfg || !map_desc || !out) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    *out = nullptr;
    if (map_desc->value_size == 0) return fail(MXCO_UNSUPPORTED, "server: StaticMap tables need a payload");
    mxco_server* s = new (std::nothrow) mxco_server();
    if (!s) return fail(MXCO_OUT_OF_MEMORY, "server: host allocation");
    s->cfg = *cfg;
    s->map_desc = *map_desc;
    if (s->map_desc.group_size == 0) s->map_desc.group_size = 4;
    // The resident kernel keeps two waves per SIMD, so by default it walks
    // one key per lane (the most keys in flight); launched ticks take the
    // tables' lane groups.
    s->res_group = s->cfg.group_size ? s->cfg.group_size : 1;
    if (!s->cfg.group_size) s->cfg.group_size = s->map_desc.group_size;
    if (!s->cfg.blocks_per_cu) s->cfg.blocks_per_cu = 64;
    if (!s->cfg.max_tick_keys) s->cfg.max_tick_keys = 1ull << 24;
    (void)hipGetDevice(&s->device);
    s->tables_cap = 4096;
    const int occ = serve_resident_occupancy(s->map_desc.key_size, s->map_desc.value_size, s->res_group);
    s->resident = s->cfg.tick_mode == 0 && env_int("MXCO_SERVE_RESIDENT", 1) != 0 && occ > 0;
    hipError_t e = hipMalloc((void**)&s->tables_dev, s->tables_cap * sizeof(BatchTable));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->map_stream, hipStreamNonBlocking);
    for (int w = 0; w < kMmWorkers && e == hipSuccess; ++w) {
        e = hipStreamCreateWithFlags(&s->mm_streams[w], hipStreamNonBlocking);
        if (e == hipSuccess && !(s->mm_ws[w] = mm_batch_ws_create())) e = hipErrorOutOfMemory;
    }
    // multimap tick buffers sized up front: no allocation on the request path
    for (int w = 0; w < kMmWorkers && e == hipSuccess; ++w)
        if (mm_batch_ws_reserve(s->mm_ws[w], kMaxTickRequests, kMaxMmTickKeys) != MXCO_OK) e = hipErrorOutOfMemory;
    for (int i = 0; i < kTickSlots && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&s->ticks[i].done, hipEventDisableTiming);
    // Resident mode (the default) for the layouts with a resident kernel;
    // tick_mode 1 or MXCO_SERVE_RESIDENT=0 keeps one launch pair per tick.
    if (s->resident && e == hipSuccess) {
        e = hipHostMalloc((void**)&s->mb, sizeof(ServeMailbox), hipHostMallocCoherent);
        if (e == hipSuccess) {
            memset((void*)s->mb, 0, sizeof(ServeMailbox));
            e = hipMalloc((void**)&s->ctl, sizeof(ServeCtl));
        }
        // a queue of its own: the running kernel never holds back the work of
        // another stream; work on the legacy default stream waits for the
        // kernel's idle exit
        if (e == hipSuccess) e = queue_stream(&s->res_stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->res_done, hipEventDisableTiming);
        int khz = 0;
        if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device);
        // The leader and one 512-thread worker per CU, which leaves room on
        // every SIMD for the multimap kernels and copies on other streams;
        // resident_blocks_per_cu > 1 asks for more workers where they fit.
        const unsigned per_cu = std::max(1u, std::min<unsigned>(s->cfg.resident_blocks_per_cu, (unsigned)occ));
        s->res_grid = 1 + (unsigned)cu_count() * per_cu;
        const double ticks_per_us = khz > 0 ? khz / 1000.0 : 100.0;
        s->res_ticks_per_us = ticks_per_us;
        s->res_idle = (uint64_t)(ticks_per_us * env_int("MXCO_SERVE_IDLE_US", 300));
        s->res_life = (uint64_t)(ticks_per_us * 50000.0);  // relaunched every 50 ms under load
        // a tick open for 5 s is given up (MXCO_SERVE_HARD_US shortens it, for tests of that path)
        s->res_hard = (uint64_t)(ticks_per_us * env_int("MXCO_SERVE_HARD_US", 5000000));
    }
    if (e != hipSuccess) {
        (void)mxco_server_destroy(s);
        return hip_fail(e, "server: create");
    }
    // every tick slot sized for a full tick up front: no allocation on the request path
    const size_t full_bytes =
        s->resident ? kMaxTickRequests * sizeof(ServeReq) + 16
                    : tick_layout(kMaxTickRequests, tick_chunk_bound(s->cfg.max_tick_keys, kMaxTickRequests),
                                  kMaxTickRequests)
                          .bytes;
    for (Tick& t : s->ticks)
        if (mxco_status st = grow_tick(t, full_bytes, 2 * kMaxTickRequests + 2, s->resident)) {
            (void)mxco_server_destroy(s);
            return st;
        }
    s->dispatcher = std::thread(dispatcher_loop, s);
    s->completer = std::thread(completer_loop, s);
    for (int w = 0; w < kMmWorkers; ++w) s->mm_workers[w] = std::thread(mm_loop, s, w);
    *out = s;
    return ok();
}

MXCO_API mxco_status mxco_server_destroy(mxco_server* s)
{
    if (!s) return ok();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_pending.notify_all();
    s->cv_done.notify_all();
    s->cv_ticks.notify_all();
    s->cv_mm.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    for (std::thread& w : s->mm_workers)
        if (w.joinable()) w.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_done.notify_all();
    if (s->completer.joinable()) s->completer.join();
    if (s->map_stream) (void)hipStreamSynchronize(s->map_stream);
    for (hipStream_t st : s->mm_streams)
        if (st) (void)hipStreamSynchronize(st);
    res_quiesce(s);
    for (Table& t : s->tables) {
        if (t.map) (void)mxco_oa_destroy(t.map);
        if (t.mm) (void)mxco_mm_destroy(t.mm);
    }
    for (Tick& t : s->ticks) {
        if (t.blob_host) (void)hipHostFree(t.blob_host);
        if (t.blob_dev) (void)hipFree(t.blob_dev);
        if (t.counters_host) (void)hipHostFree(t.counters_host);
        for (void* p : t.retired_host) (void)hipHostFree(p);
        for (void* p : t.retired_dev) (void)hipFree(p);
        if (t.done) (void)hipEventDestroy(t.done);
    }
    for (auto& kv : s->reqs) delete kv.second;
    if (s->res_stream) (void)hipStreamSynchronize(s->res_stream);
    if (s->res_done) (void)hipEventDestroy(s->res_done);
    if (s->res_stream) (void)hipStreamDestroy(s->res_stream);
    if (s->ctl) (void)hipFree(s->ctl);
    if (s->mb) (void)hipHostFree(s->mb);
    if (s->tables_dev) (void)hipFree(s->tables_dev);
    if (s->map_stream) (void)hipStreamDestroy(s->map_stream);
    for (hipStream_t st : s->mm_streams)
        if (st) (void)hipStreamDestroy(st);
    for (MmBatchWs* ws : s->mm_ws) mm_batch_ws_destroy(ws);
    delete s;
    return ok();
}

MXCO_API mxco_status mxco_server_add_map(mxco_server* s, uint64_t capacity, uint32_t* table)
{
    if (!s || !table) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    const uint64_t cap = valid_extent(capacity, s->map_desc.bucket_size, s->map_desc.probing);
    const uint32_t ss = (s->map_desc.key_size == 4 && s->map_desc.value_size == 4) ? 8 : 16;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->tables.size() >= s->tables_cap) return fail(MXCO_LENGTH_ERROR, "server: too many tables");
    if (mxco_status e = admit(s, cap * ss)) return e;
    mxco_oa* m = nullptr;
    if (mxco_status e = mxco_oa_create(&s->map_desc, capacity, s->cfg.empty_key, s->cfg.empty_value, s->map_stream, &m))
        return e;
    // the table's fill is done before any tick can reach it (the resident
    // kernel runs on a stream of its own)
    if (hipError_t e = hipStreamSynchronize(s->map_stream); e != hipSuccess) {
        (void)mxco_oa_destroy(m);
        return hip_fail(e, "server: add table");
    }
    BatchTable bt;
    void* slots = nullptr;
    uint64_t c = 0;
    (void)mxco_oa_data(m, &slots);
    (void)mxco_oa_capacity(m, &c);
    bt.slots = (uint8_t*)slots;
    bt.cap = (uint32_t)c;
    bt.nbuckets = (uint32_t)(c / s->map_desc.bucket_size);
    const uint32_t id = (uint32_t)s->tables.size();
    // a new descriptor entry: no running kernel reads this index yet (on the
    // map stream: the default stream would wait for the resident kernel)
    MXCO_HIP(hipMemcpyAsync(s->tables_dev + id, &bt, sizeof(bt), hipMemcpyHostToDevice, s->map_stream),
             "server: add table");
    MXCO_HIP(hipStreamSynchronize(s->map_stream), "server: add table");
    Table t;
    t.kind = 0;
    t.map = m;
    t.bt = bt;
    t.bytes = c * ss;
    t.live = true;
    s->tables.push_back(t);
    *table = id;
    return ok();
}

MXCO_API mxco_status mxco_server_add_multimap(mxco_server* s, uint64_t capacity, const void* empty_key,
                                              const void* empty_value, uint32_t* table)
{
    if (!s || !table) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->tables.size() >= s->tables_cap) return fail(MXCO_LENGTH_ERROR, "server: too many tables");
    uint64_t p = 64;
    while (p < 2 * capacity) p <<= 1;
    if (mxco_status e = admit(s, p * 16)) return e;
    mxco_mm* mm = nullptr;
    hipStream_t stream = s->mm_streams[s->tables.size() % kMmWorkers];  // the stream of the table's worker
    if (mxco_status e = mxco_mm_create(8, 8, capacity, empty_key, empty_value, stream, &mm)) return e;
    Table t;
    t.kind = 1;
    t.mm = mm;
    t.bytes = p * 16;
    t.live = true;
    s->tables.push_back(t);
    *table = (uint32_t)(s->tables.size() - 1);
    return ok();
}

MXCO_API mxco_status mxco_server_drop_table(mxco_server* s, uint32_t table)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    {
        // from here on submits to the table fail, so the drain below also
        // waits for every request that reached it before
        std::lock_guard<std::mutex> lk(s->mu);
        if (table >= s->tables.size() || !s->tables[table].live)
            return fail(MXCO_INVALID_ARGUMENT, "server: unknown table");
        s->tables[table].live = false;
    }
    (void)mxco_server_drain(s);
    std::lock_guard<std::mutex> lk(s->mu);
    res_quiesce(s);  // hipFree would wait for the resident kernel's idle exit
    Table& t = s->tables[table];
    if (t.map) (void)mxco_oa_destroy(t.map);
    if (t.mm) (void)mxco_mm_destroy(t.mm);
    t.map = nullptr;
    t.mm = nullptr;
    t.live = false;
    return ok();
}

MXCO_API mxco_status mxco_server_submit(mxco_server* s, const mxco_request* req, uint64_t* ticket)
{
    if (!s || !req || !ticket) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    Req* r = new (std::nothrow) Req();
    if (!r) return fail(MXCO_OUT_OF_MEMORY, "server: host allocation");
    r->r = *req;
    r->t_submit = now_us();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (mxco_status e = validate(s, req)) {
            delete r;
            return e;
        }
        r->bt = s->tables[req->table].bt;
        r->ticket = s->next_ticket++;
        s->reqs[r->ticket] = r;
        if (r->r.n == 0) {
            finish(s, r, 0, MXCO_OK);
        } else {
            s->pending.push_back(r);
        }
    }
    s->cv_pending.notify_one();
    *ticket = r->ticket;
    return ok();
}

MXCO_API mxco_status mxco_server_wait(mxco_server* s, uint64_t ticket, uint64_t* result, double* latency_us)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(MXCO_INVALID_ARGUMENT, "server: unknown ticket");
    Req* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    if (result) *result = r->result;
    if (latency_us) *latency_us = r->t_done - r->t_submit;
    const mxco_status st = r->status;
    s->reqs.erase(it);
    delete r;
    if (st != MXCO_OK) return fail(st, "server: request failed: " + s->worker_error);
    return ok();
}

MXCO_API mxco_status mxco_server_drain(mxco_server* s)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (!s->pending.empty() || !s->inflight.empty()) return false;
        for (const std::deque<Req*>& q : s->mm_pending)
            if (!q.empty()) return false;
        for (auto& kv : s->reqs)
            if (!kv.second->done) return false;
        for (Tick& t : s->ticks)
            if (t.busy) return false;
        return true;
    });
    return ok();
}

MXCO_API mxco_status mxco_server_replay(mxco_server* s, const mxco_request* reqs, const double* arrival_us,
                                        uint32_t n, uint64_t* results, double* latency_us)
{
    if (!s || (n && (!reqs || !arrival_us))) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::vector<uint64_t> tickets(n);
    std::vector<double> late(n);  // submit time - arrival time
    mxco_status first_err = MXCO_OK;
    std::string err_msg;
    const double t0 = now_us();
    for (uint32_t i = 0; i < n; ++i) {
        const double due = t0 + arrival_us[i];
        double t = now_us();
        while (t < due) {  // open loop: submit at the scheduled arrival time
            if (due - t > 200) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(due - t - 100)));
            t = now_us();
        }
        late[i] = t - due;
        if (mxco_status e = mxco_server_submit(s, &reqs[i], &tickets[i])) {
            tickets[i] = 0;
            if (first_err == MXCO_OK) {
                first_err = e;
                err_msg = mxco_last_error();
            }
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t res = 0;
        double lat = 0;
        if (tickets[i]) {
            mxco_status e = mxco_server_wait(s, tickets[i], &res, &lat);
            if (e != MXCO_OK && first_err == MXCO_OK) {
                first_err = e;
                err_msg = mxco_last_error();
            }
        }
        if (results) results[i] = res;
        if (latency_us) latency_us[i] = lat + late[i];  // from the arrival time
    }
    if (first_err != MXCO_OK) return fail(first_err, err_msg);
    return ok();
}

MXCO_API mxco_status mxco_server_stats(mxco_server* s, uint64_t* ticks, uint64_t* tick_requests)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticks) *ticks = s->n_ticks;
    if (tick_requests) *tick_requests = s->n_tick_requests;
    return ok();
}

MXCO_API mxco_status mxco_server_resident_stats(mxco_server* s, uint32_t* resident, uint64_t* launches,
                                                uint32_t* grid)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (resident) *resident = s->resident;
    if (launches) *launches = s->res_launches;
    if (grid) *grid = s->res_grid;
    return ok();
}

MXCO_API mxco_status mxco_server_resident_trace(mxco_server* s, uint64_t* out, uint32_t max_ticks, uint32_t* n,
                                                double* ticks_per_us)
{
    if (!s || (max_ticks && !out)) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    if (mxco_status e = mxco_server_drain(s)) return e;
    std::lock_guard<std::mutex> lk(s->mu);
    if (n) *n = 0;
    if (!s->resident) return ok();
    res_quiesce(s);
    const uint32_t k = (uint32_t)std::min<uint64_t>({(uint64_t)max_ticks, (uint64_t)kServeTrace, s->res_posted});
    uint64_t* all = nullptr;
    MXCO_HIP(hipHostMalloc((void**)&all, sizeof(uint64_t) * 4 * kServeTrace, 0), "server: trace");
    hipError_t e = hipMemcpyAsync(all, s->ctl->trace, sizeof(uint64_t) * 4 * kServeTrace, hipMemcpyDeviceToHost,
                                  s->map_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->map_stream);
    if (e == hipSuccess) {
        // the last k ticks, oldest first
        for (uint32_t i = 0; i < k; ++i) {
            const uint64_t seq = s->res_posted - k + 1 + i;
            memcpy(out + 4 * i, all + 4 * (seq % kServeTrace), 4 * sizeof(uint64_t));
        }
        if (n) *n = k;
        if (ticks_per_us) *ticks_per_us = s->res_ticks_per_us;
    }
    (void)hipHostFree(all);
    if (e != hipSuccess) return hip_fail(e, "server: trace");
    return ok();
}

MXCO_API mxco_status mxco_server_table(mxco_server* s, uint32_t table, mxco_oa** map, mxco_mm** multimap)
{
    if (!s) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (table >= s->tables.size() || !s->tables[table].live) return fail(MXCO_INVALID_ARGUMENT, "server: unknown table");
    if (map) *map = s->tables[table].map;
    if (multimap) *multimap = s->tables[table].mm;
    return ok();
}

// The same open-loop replay, served one request at a time with the plain
// synchronous calls (the unbatched baseline): request i starts at
// max(its arrival, the previous completion).
MXCO_API mxco_status mxco_server_replay_plain(mxco_server* s, const mxco_request* reqs, const double* arrival_us,
                                              uint32_t n, uint64_t* results, double* latency_us)
{
    if (!s || (n && (!reqs || !arrival_us))) return fail(MXCO_INVALID_ARGUMENT, "server: null argument");
    if (mxco_status e = mxco_server_drain(s)) return e;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        res_quiesce(s);  // the baseline runs without the resident kernel holding CUs
    }
    hipStream_t st = s->map_stream;
    const double t0 = now_us();
    for (uint32_t i = 0; i < n; ++i) {
        const double due = t0 + arrival_us[i];
        double t = now_us();
        while (t < due) {
            if (due - t > 200) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(due - t - 100)));
            t = now_us();
        }
        const mxco_request& r = reqs[i];
        if (r.table >= s->tables.size() || !s->tables[r.table].live)
            return fail(MXCO_INVALID_ARGUMENT, "server: unknown table");
        Table& tb = s->tables[r.table];
        uint64_t res = 0;
        mxco_status e = MXCO_OK;
        switch (r.op) {
        case MXCO_REQ_INSERT: e = mxco_oa_insert(tb.map, r.in, nullptr, r.n, st, 1, &res); break;
        case MXCO_REQ_FIND: e = mxco_oa_find(tb.map, r.in, nullptr, r.out, r.n, st, 1); break;
        case MXCO_REQ_CONTAINS: e = mxco_oa_contains(tb.map, r.in, nullptr, (uint8_t*)r.out, r.n, st, 1); break;
        case MXCO_REQ_MM_INSERT: e = mxco_mm_insert(tb.mm, r.in, nullptr, r.n, st, &res); break;
        case MXCO_REQ_MM_FIND: e = mxco_mm_find(tb.mm, r.in, nullptr, r.out, r.n, st); break;
        case MXCO_REQ_MM_COUNT: e = mxco_mm_count(tb.mm, r.in, r.n, st, &res); break;
        case MXCO_REQ_MM_RETRIEVE:
            e = mxco_mm_retrieve(tb.mm, r.in, r.n, r.out, r.out2, r.out_capacity, st, &res);
            break;
        default: e = fail(MXCO_INVALID_ARGUMENT, "server: unknown request");
        }
        if (e != MXCO_OK) return e;
        if (results) results[i] = res;
        if (latency_us) latency_us[i] = now_us() - due;
    }
    return ok();
}

}  // extern "C"
