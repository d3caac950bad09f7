// Resident serving kernel: one long-running launch that takes the serving
// engine's StaticMap ticks from a queue, so that a tick costs no kernel
// launch, no copy and no read-back.
//
// Queue.  The host writes a tick's request list (per request: table,
// operation, device input and output, key count, keys per chunk and the
// index of its first chunk; inserts first) into host-coherent pinned memory,
// fills the tick's descriptor in the mailbox and publishes it by a release
// store of `posted`.  Ticks are numbered 1, 2, ... and use descriptor slot
// seq % kServeSlots; the host posts a tick only when the tick that used its
// slot before has finished.  Only the leader reads host memory (a GPU read
// of host memory is a PCIe round trip, and few can be in flight): the
// descriptor, and the request list with coalesced 16-byte loads.  The
// kernel writes each tick's insert counters and done word into the mailbox,
// where the host polls them.  The kernel depends on no other stream.
//
// Workgroups.  Workgroup 0 leads: it polls the mailbox, and when the open
// tick has finished and the next one is posted, it copies that tick's
// descriptor and request list into ServeCtl, zeroes the tick's insert
// counters and opens the tick as a new epoch by a release store of the claim
// word [epoch:16 | chunks:24 | next chunk:24].  The other workgroups (one of
// 512 threads per CU, leaving room on every SIMD for the kernels of other
// streams) claim chunks with a fetch-and-add on that word: the value
// returned names the epoch and its chunk count, so a claim is valid by
// itself (a claim past the count takes nothing; there is at most one such
// claim per workgroup and epoch, so the next-chunk field never overflows).
// While the tick has more chunks left than there are workers, a worker
// issues its next claim when a chunk starts, so the claim's round trip is
// hidden by the chunk's work.  A worker finds a chunk's request by binary
// search over the requests' first-chunk indices, held in its LDS for the
// epoch with the requests themselves (up to kServeReqCache; each carries its
// table's slot array, capacity and bucket count).  Insert chunks, and
// read chunks of fewer than kServeStagedMin keys, run the probe state
// machine of the launched path (oa_process) on the chunk's keys.  Larger read
// chunks run the bulk Find's wave-staged loop (wave_staged_find): each
// wavefront takes 128-key pieces of the chunk, with the keys copied into LDS
// when a piece opens and the outputs written back as whole lines.  Insert
// chunks come first.  A read of a table with an insert in the same tick waits
// until every insert chunk of the tick has been processed and released, so it
// sees the tick's inserts, as it does after the launched path's insert kernel;
// other reads start at once.
//
// Memory ordering (on gfx950 an agent-scope release writes back this XCD's
// L2 and an acquire invalidates it, so each is done once per workgroup and
// tick, not per chunk; polls and claims are relaxed).  A workgroup releases
// its insert chunks when it leaves the insert phase, and its outputs at
// system scope when it finds no chunk left; it then adds its chunks to the
// done count and leaves the writer count.  The workgroup whose leave brings
// the writer count to zero with every chunk done finishes the tick: it
// copies the insert counters to the host, stores the tick's number into the
// mailbox's done word of the slot and marks the epoch finished for the
// leader.  A worker acquires at system scope when it first claims a chunk of
// a new epoch, so request inputs written before the tick was posted are
// read fresh.
//
// Lifetime.  Every wait ends: workers wait only for chunks that running
// workgroups have claimed, or for the leader's exit flag; the leader sets
// that flag between ticks when the host sets the mailbox's stop word, when
// no tick has come for `idle` clock ticks, or when the launch is older than
// `life` (or has opened kServeMaxEpochs ticks), and in the middle of a tick
// only when the tick has been open for `hard` clock ticks (it then reports
// the tick as aborted, by its number).  A worker that saw the exit flag
// when it claimed ahead drops that claim.  The host relaunches the kernel
// whenever a posted tick is unfinished and the kernel has ended.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

#include "batch_kernels.h"

namespace mxco {

constexpr int kServeSlots = 4;                 // ticks in flight (one blob slot each)
constexpr unsigned kServeThreads = 512;        // threads per workgroup: two waves per SIMD, one worker per CU
constexpr uint32_t kServeMaxItems = 4096;      // requests per tick
constexpr uint32_t kServeMaxChunks = 1u << 23; // chunks per tick (the claim word's 24-bit fields)
constexpr uint32_t kServeMaxEpochs = 60000;    // ticks per launch (the claim word's 16-bit epoch)
constexpr uint32_t kServeTrace = 1024;         // ticks kept in the timing trace
// ServeReq::op flag of a read whose table has an insert in the same tick:
// its chunks wait for the tick's inserts (other reads start at once).
constexpr uint32_t kServeAfterInsert = 0x80000000u;
// Requests per tick that a worker keeps in LDS (larger ticks read them from
// device memory).
constexpr uint32_t kServeReqCache = 256;
// Read chunks of at least this many keys run the wave-staged loop: every
// wavefront of the worker gets at least one 128-key piece.
constexpr uint32_t kServeStagedMin = (kServeThreads / 64) * kWaveChunk;

// One request of a tick; its chunks are [first, first + ceil(count / csz)).
struct ServeReq {
    const uint8_t* in;    // device input (pairs or keys)
    uint8_t* out;         // device output (reads)
    uint64_t count;       // keys
    uint8_t* slots;       // the table: slot array, capacity and bucket count
    uint32_t cap, nbuckets;
    uint32_t op;          // MXCO_REQ_INSERT, MXCO_REQ_FIND or MXCO_REQ_CONTAINS (| kServeAfterInsert)
    uint32_t csz;         // keys per chunk
    uint32_t first, pad;  // the request's first chunk in the tick
};

struct ServeTickDesc {
    const ServeReq* reqs;               // host-coherent request list, inserts first
    unsigned long long* counters_host;  // host-coherent: 2 per insert request, written when the tick finishes
    uint64_t seq;                       // tick number (1, 2, ...)
    uint32_t n_reqs, n_insert_items;
    uint32_t n_ins_chunks, n_rd_chunks;
};

// Host-coherent memory shared by the host and the resident kernel.
struct ServeMailbox {
    uint64_t posted;                  // host: seq of the last tick posted
    uint64_t stop;                    // host: exit at the next moment between ticks
    uint64_t aborted;                 // kernel: number of an unfinished tick the hard deadline ended
    uint64_t pad;
    uint64_t done[kServeSlots];       // kernel: seq of the last finished tick of each slot
    ServeTickDesc desc[kServeSlots];  // host: the tick posted into slot seq % kServeSlots
};

// Device memory of the resident kernel; the header (up to `cur`) is zeroed
// before each launch.
// The words that atomics hit sit on lines of their own, so the claims, the
// counts and the flags do not queue behind each other.
struct ServeCtl {
    __attribute__((aligned(128))) uint64_t claim;  // [epoch:16 | chunks:24 | next chunk:24]
    // [chunks of the epoch processed:32 | workgroups holding processed,
    // unreleased chunks of the epoch:32]; a workgroup adds its chunks and
    // leaves in one atomic
    __attribute__((aligned(128))) uint64_t done_writers;
    uint32_t ins_done;  // insert chunks of the epoch processed and released
    __attribute__((aligned(128))) uint32_t finished;  // last epoch whose results are published
    uint32_t exit;
    __attribute__((aligned(128))) ServeTickDesc cur;  // descriptor of the open epoch
    // Per epoch parity: the insert counters and the request list.  The leader
    // stages the next tick into the other parity while the open one runs.
    unsigned long long counters[2][2 * kServeMaxItems];
    __attribute__((aligned(16))) ServeReq reqs[2][kServeMaxItems];
    // Timing trace of the last kServeTrace ticks, by seq % kServeTrace (wall
    // clock): the leader began staging the tick, opened it, the finisher
    // published it; [insert chunks:32 | read chunks:32].
    uint64_t trace[kServeTrace][4];
};

constexpr size_t kServeCtlHeader = offsetof(ServeCtl, cur);

__device__ __forceinline__ uint32_t claim_next(uint64_t v) { return (uint32_t)v & 0xffffffu; }
__device__ __forceinline__ uint32_t claim_count(uint64_t v) { return (uint32_t)(v >> 24) & 0xffffffu; }
__device__ __forceinline__ uint32_t claim_epoch(uint64_t v) { return (uint32_t)(v >> 48); }

template <typename T>
__device__ __forceinline__ T srv_peek(T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T srv_peek_sys(T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Plain-old-data copy out of host-coherent memory, 8 bytes at a time.
template <typename T>
__device__ __forceinline__ T srv_copy_sys(const T* p)
{
    static_assert(sizeof(T) % 8 == 0, "8-byte words");
    T v;
    uint64_t* d = reinterpret_cast<uint64_t*>(&v);
    uint64_t* s = const_cast<uint64_t*>(reinterpret_cast<const uint64_t*>(p));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 8); ++i) d[i] = srv_peek_sys(s + i);
    return v;
}

// Workgroup 0: stages the next posted tick (descriptor and request list)
// while the open one runs, opens it the moment the open one finishes, and
// decides the exit.
__device__ __forceinline__ void serve_lead(ServeMailbox* mb, ServeCtl* ctl, uint64_t seq_base, uint64_t idle,
                                           uint64_t life, uint64_t hard)
{
    __shared__ uint32_t s_act, s_par;
    __shared__ ServeTickDesc s_d;
    uint32_t e = 0;       // open epoch (thread 0)
    bool staged = false;  // epoch e + 1 is staged (thread 0)
    ServeTickDesc next{};  // its descriptor (thread 0)
    const uint64_t t0 = wall_clock64();
    uint64_t last = t0, opened = t0;  // last moment a tick was open; when the open tick was published
    uint64_t open_seq = 0;            // number of the open tick (thread 0)
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t act = 0;  // 0 wait, 1 stage, 2 open, 3 exit, 4 abort
            const uint64_t now = wall_clock64();
            // relaxed polls: an acquire here would invalidate this XCD's L2 on every poll
            const bool idle_now = srv_peek(&ctl->finished) == e;
            if (!idle_now) last = now;
            if (!staged && e + 1 < kServeMaxEpochs && srv_peek_sys(&mb->posted) > seq_base + e) {
                act = 1;
            } else if (idle_now) {
                if (staged) act = 2;
                else if (srv_peek_sys(&mb->stop) || e + 1 >= kServeMaxEpochs) act = 3;
                else if (now - last > idle || now - t0 > life) act = 3;
            } else if (now - opened > hard) {
                act = 4;
            }
            if (act == 1) {
                // the host's descriptor and list; the parity's counters and list
                // were last read by the finisher and workers of epoch e - 1
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                next = srv_copy_sys(&mb->desc[(seq_base + e + 1) % kServeSlots]);
                s_d = next;
                s_par = (e + 1) & 1;
                ctl->trace[next.seq % kServeTrace][0] = now;
            }
            s_act = act;
        }
        __syncthreads();
        const uint32_t act = s_act;
        if (act == 1) {
            // the request list: coalesced 16-byte loads from host memory
            const ServeTickDesc d = s_d;
            const uint32_t par = s_par;  // thread 0 alone counts the epochs
            const uint32_t n16 = (d.n_reqs * (uint32_t)sizeof(ServeReq) + 15) / 16;
            const uint4* src = reinterpret_cast<const uint4*>(d.reqs);
            uint4* dst = reinterpret_cast<uint4*>(ctl->reqs[par]);
            for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
            for (uint32_t i = threadIdx.x; i < 2 * d.n_insert_items; i += blockDim.x) ctl->counters[par][i] = 0;
            __syncthreads();
            if (threadIdx.x == 0) staged = true;
            continue;
        }
        __syncthreads();  // every wave has read s_act before thread 0 writes it again
        if (act == 2) {
            if (threadIdx.x == 0) {
                ++e;
                staged = false;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the finisher's reads before the resets
                ctl->cur = next;
                open_seq = next.seq;
                ctl->done_writers = 0;
                ctl->ins_done = 0;
                last = opened = wall_clock64();
                ctl->trace[next.seq % kServeTrace][1] = opened;
                ctl->trace[next.seq % kServeTrace][3] = ((uint64_t)next.n_ins_chunks << 32) | next.n_rd_chunks;
                // the staged list, descriptor, counters and resets reach every XCD before the epoch opens
                const uint64_t claim =
                    ((uint64_t)e << 48) | ((uint64_t)(next.n_ins_chunks + next.n_rd_chunks) << 24);
                __hip_atomic_store(&ctl->claim, claim, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            continue;
        }
        if (act >= 3) {
            if (threadIdx.x == 0) {
                if (act == 4)
                    __hip_atomic_store(&mb->aborted, open_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&ctl->exit, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// Thread 0 of a worker: the tick's insert chunks it processed are released
// and counted (reads of the tick wait for all of them).
__device__ __forceinline__ void serve_release_inserts(ServeCtl* ctl, uint32_t& ins_mine)
{
    if (ins_mine) {
        __hip_atomic_fetch_add(&ctl->ins_done, ins_mine, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        ins_mine = 0;
    }
}

// Thread 0 of a worker that processed chunks of the epoch and found none
// left: releases its outputs at system scope, then adds its chunks to the
// done count and leaves the writer count in one atomic; the last writer out,
// with every chunk done, finishes the tick.
__device__ __forceinline__ void serve_leave(ServeMailbox* mb, ServeCtl* ctl, const ServeTickDesc& d, uint32_t epoch,
                                            uint32_t& ins_mine, uint32_t& done_mine)
{
    // one release covers the outputs, the insert count and the done / writer word below
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (ins_mine) __hip_atomic_fetch_add(&ctl->ins_done, ins_mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ins_mine = 0;
    // + done_mine chunks, - 1 writer (this workgroup holds at least one, so no borrow crosses the halves)
    const uint64_t old = __hip_atomic_fetch_add(&ctl->done_writers, ((uint64_t)done_mine << 32) - 1,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t left = (uint32_t)old - 1;
    const uint32_t done = (uint32_t)(old >> 32) + done_mine;
    done_mine = 0;
    if (left != 0 || done != d.n_ins_chunks + d.n_rd_chunks) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    for (uint32_t i = 0; i < 2 * d.n_insert_items; ++i)
        __hip_atomic_store(d.counters_host + i, srv_peek(&ctl->counters[epoch & 1][i]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    ctl->trace[d.seq % kServeTrace][2] = wall_clock64();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __hip_atomic_store(&mb->done[d.seq % kServeSlots], d.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&ctl->finished, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 128-key pieces of one read chunk (no stencil), held by value: wavefront w
// takes pieces w, w + waves, ...
template <int KS, int OUTB>
struct ServePieces {
    static constexpr bool kOneTable = true;
    uint8_t* table;
    const uint8_t* in;
    uint8_t* out;
    uint64_t n, next, step;
    uint32_t cap, nbuckets;
    __device__ __forceinline__ WsChunk open()
    {
        WsChunk c;
        const uint64_t base = next * kWaveChunk;
        next += step;
        c.table = table;
        c.cap = cap;
        c.nbuckets = nbuckets;
        c.len = base < n ? (uint32_t)min((uint64_t)kWaveChunk, n - base) : 0u;
        c.in = in + base * KS;
        c.out = out + base * OUTB;
        c.stencil = nullptr;
        return c;
    }
};

template <int KS, int VS, int G>
__device__ __forceinline__ void serve_work(const OaArgs& base, ServeMailbox* mb, ServeCtl* ctl)
{
    __shared__ uint32_t s_first[kServeMaxItems];  // first chunk of each request of the epoch
    __shared__ ServeReq s_reqs[kServeReqCache];   // the epoch's requests, when they fit
    // per wavefront: output staging and key buffers of its two open pieces
    __shared__ __align__(16) uint8_t s_out[kServeThreads / 64][2][kWaveChunk * 8];
    __shared__ __align__(16) uint8_t s_keys[kServeThreads / 64][2][kWaveChunk * 8];
    __shared__ ServeReq s_req;
    __shared__ uint32_t s_state, s_idx, s_nreq, s_nins, s_fresh, s_epoch;
    // thread 0's state
    ServeTickDesc d{};
    uint32_t my_e = 0, ins_mine = 0, done_mine = 0, backoff = 1;
    uint32_t dead_e = 0xffffffffu;  // epoch in which this workgroup's claim came back empty
    uint64_t ahead = 0;             // claim issued while the previous chunk ran
    uint32_t exit_ahead = 0;        // the exit flag, read with it
    bool joined = false, ins_seen = false, have_ahead = false;
    const uint32_t gpb = blockDim.x / G;
    const uint32_t gj = threadIdx.x / G;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t state = 0, idx = 0, fresh = 0;
            uint64_t old = ahead;
            bool got = have_ahead && !exit_ahead;  // after the exit flag: no further chunk
            have_ahead = false;
            for (;;) {
                if (got) {
                    got = false;
                    if (claim_next(old) < claim_count(old)) {
                        idx = claim_next(old);
                        if (claim_epoch(old) != my_e) {
                            // a chunk of a new tick (this workgroup holds none of an older one):
                            // inputs and the request list written before it opened
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                            d = ctl->cur;
                            my_e = claim_epoch(old);
                            ins_seen = false;
                            fresh = 1;
                        }
                        state = 1;
                        backoff = 1;
                        break;
                    }
                    // the tick has no chunk left: this workgroup claims no more in it
                    dead_e = claim_epoch(old);
                }
                if (joined) {
                    serve_leave(mb, ctl, d, my_e, ins_mine, done_mine);
                    joined = false;
                }
                if (srv_peek(&ctl->exit)) {
                    state = 2;
                    break;
                }
                const uint64_t v = srv_peek(&ctl->claim);
                if (claim_epoch(v) != dead_e && claim_next(v) < claim_count(v)) {
                    old = __hip_atomic_fetch_add(&ctl->claim, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    got = true;
                    continue;
                }
                for (uint32_t k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(2);
                backoff = backoff < 8 ? 2 * backoff : 8;
            }
            if (state == 1) {
                // The next chunk's claim, issued now so that it returns while this
                // chunk runs, when the tick has more chunks left than workers (in a
                // smaller tick it would take a chunk an idle worker could start at
                // once).  A claim past the tick's chunks is this workgroup's one
                // empty claim of the tick.
                if (claim_count(old) - claim_next(old) > gridDim.x) {
                    ahead = __hip_atomic_fetch_add(&ctl->claim, (uint64_t)1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    exit_ahead = srv_peek(&ctl->exit);
                    have_ahead = true;
                }
                if (!joined) {
                    __hip_atomic_fetch_add(&ctl->done_writers, (uint64_t)1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    joined = true;
                }
                // a read chunk: this workgroup's inserts of the tick are released for
                // the reads that wait for them
                if (idx >= d.n_ins_chunks) serve_release_inserts(ctl, ins_mine);
                s_idx = idx;
                s_nreq = d.n_reqs;
                s_nins = d.n_ins_chunks;
                s_fresh = fresh;
                s_epoch = my_e;
            }
            s_state = state;
        }
        __syncthreads();
        if (s_state == 2) return;
        const uint32_t nreq = s_nreq;
        const ServeReq* reqs = ctl->reqs[s_epoch & 1];
        if (s_fresh) {
            for (uint32_t i = threadIdx.x; i < nreq; i += blockDim.x) s_first[i] = reqs[i].first;
            if (nreq <= kServeReqCache) {
                constexpr uint32_t W = sizeof(ServeReq) / 8;
                const uint64_t* src = reinterpret_cast<const uint64_t*>(reqs);
                uint64_t* dst = reinterpret_cast<uint64_t*>(s_reqs);
                for (uint32_t i = threadIdx.x; i < nreq * W; i += blockDim.x) dst[i] = src[i];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            // the chunk's request: the last one whose first chunk is at most idx
            const uint32_t idx = s_idx;
            uint32_t lo = 0, hi = nreq;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_first[mid] <= idx) lo = mid;
                else hi = mid;
            }
            ServeReq rq = nreq <= kServeReqCache ? s_reqs[lo] : reqs[lo];
            if ((rq.op & kServeAfterInsert) && !ins_seen) {
                // a read of a table with an insert in this tick: after all of the tick's inserts
                while (srv_peek(&ctl->ins_done) < d.n_ins_chunks && !srv_peek(&ctl->exit))
                    __builtin_amdgcn_s_sleep(2);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ins_seen = true;
            }
            rq.op &= ~kServeAfterInsert;
            rq.first = lo;  // from here on: the request's index (its insert counter pair)
            s_req = rq;
        }
        __syncthreads();
        const uint32_t idx = s_idx;
        const ServeReq rq = s_req;
        const uint32_t item = rq.first;
        const bool ins = idx < s_nins;
        const uint64_t start = (uint64_t)(idx - s_first[item]) * rq.csz;
        const uint64_t end = min(start + rq.csz, rq.count);
        OaArgs a = base;
        a.table = rq.slots;
        a.cap = rq.cap;
        a.nbuckets = rq.nbuckets;
        a.in = rq.in;
        a.out = rq.out;
        a.n = rq.count;
        unsigned long long c0 = 0, c1 = 0;
        if (ins) {
            oa_process<KS, VS, G, OP_INSERT>(a, start + gj, end, gpb, c0, c1);
            flush_counters<OP_INSERT>(ctl->counters[s_epoch & 1] + 2ull * item, c0, c1);
        } else if (end - start >= kServeStagedMin) {
            // wavefront w takes the chunk's 128-key pieces w, w + waves, ...
            constexpr int OUTF = VS ? VS : KS;
            const uint32_t wib = threadIdx.x >> 6, waves = blockDim.x >> 6;
            a.stencil = nullptr;
            const uint64_t n = end - start;
            const uint8_t* in = rq.in + start * KS;
            if (rq.op == MXCO_REQ_CONTAINS) {
                ServePieces<KS, 1> src{rq.slots, in, rq.out + start, n, wib, waves, rq.cap, rq.nbuckets};
                wave_staged_find<KS, VS, G, OP_CONTAINS>(a, src, s_out[wib][0], s_out[wib][1], s_keys[wib][0],
                                                        s_keys[wib][1]);
            } else {
                ServePieces<KS, OUTF> src{rq.slots, in, rq.out + start * OUTF, n, wib, waves, rq.cap, rq.nbuckets};
                wave_staged_find<KS, VS, G, OP_FIND>(a, src, s_out[wib][0], s_out[wib][1], s_keys[wib][0],
                                                    s_keys[wib][1]);
            }
        } else if (rq.op == MXCO_REQ_CONTAINS) {
            oa_process<KS, VS, G, OP_CONTAINS>(a, start + gj, end, gpb, c0, c1);
        } else {
            oa_process<KS, VS, G, OP_FIND>(a, start + gj, end, gpb, c0, c1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            ins_mine += ins;
            ++done_mine;
        }
    }
}

template <int KS, int VS, int G>
__global__ __launch_bounds__(kServeThreads) void serve_resident_kernel(OaArgs base, ServeMailbox* mb, ServeCtl* ctl,
                                                                       uint64_t seq_base, uint64_t idle, uint64_t life,
                                                                       uint64_t hard)
{
    if (blockIdx.x == 0) serve_lead(mb, ctl, seq_base, idle, life, hard);
    else serve_work<KS, VS, G>(base, mb, ctl);
}

// Launches the resident kernel for the map layout (defined in
// serve_resident.hip); hipErrorInvalidValue for a layout without one.
hipError_t launch_serve_resident(int ks, int vs, int g, const OaArgs& a, ServeMailbox* mb,
                                 ServeCtl* ctl, uint64_t seq_base, uint64_t idle, uint64_t life, uint64_t hard,
                                 unsigned grid, hipStream_t s);

// Workgroups of the resident kernel that fit one CU at once for the layout (0 = none).
int serve_resident_occupancy(int ks, int vs, int g);

}  // namespace mxco
