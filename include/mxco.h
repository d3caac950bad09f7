/*
 * mxco.h -- C ABI of libmxco.so, the MI355X (gfx950) implementation of the
 * ops-collections concurrent-container bulk operations.
 *
 * Every entry point takes plain pointers, sizes and a hipStream_t.  All key,
 * value, stencil and output buffers are caller-owned device pointers; the
 * library never copies user data between host and device.  Each entry point
 * cites the reference interface it replaces (file:line in ops-collections).
 *
 * Error model (reference: include/utility/kernel_launch_utils.h:19-29 and
 * the throwing sites listed per function): every call returns an
 * mxco_status.  The C++ headers in include/mxco/ turn the codes back into the
 * exception types the reference throws (std::invalid_argument,
 * std::length_error, std::bad_alloc, std::runtime_error, std::logic_error).
 * mxco_last_error() returns the calling thread's last message.
 *
 * Threading: calls on one object must be serialised by the caller, exactly
 * as in the reference (per-object scratch counters,
 * include/detail/open_addressing/open_addressing_impl.h:548-558).
 */
#ifndef MXCO_H
#define MXCO_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MXCO_API __attribute__((visibility("default")))

typedef enum mxco_status {
    MXCO_OK = 0,
    MXCO_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    MXCO_LENGTH_ERROR = 2,     /* std::length_error (batch beyond the index range) */
    MXCO_OUT_OF_MEMORY = 3,    /* std::bad_alloc (device allocation failed) */
    MXCO_RUNTIME_ERROR = 4,    /* std::runtime_error (HIP runtime failure) */
    MXCO_LOGIC_ERROR = 5,      /* std::logic_error (moved-from / uninitialised object) */
    MXCO_UNSUPPORTED = 6,      /* configuration this build does not provide */
    MXCO_OVERFLOW_ERROR = 7    /* std::overflow_error (a count that may not fit 64 bits) */
} mxco_status;

/* Last error message of the calling thread ("" after success). */
MXCO_API const char* mxco_last_error(void);
MXCO_API const char* mxco_status_string(mxco_status status);
/* Library version, e.g. "0.1.0". */
MXCO_API const char* mxco_version(void);

/* ---------------------------------------------------------------------- */
/* Shared enums                                                           */
/* ---------------------------------------------------------------------- */

/* ProbingScheme (include/probing_scheme.h:37-122). */
typedef enum mxco_probing { MXCO_PROBING_LINEAR = 0, MXCO_PROBING_DOUBLE = 1 } mxco_probing;

/* Hasher families (include/hash_functions.h:17-30).  Seeds are fixed as in
 * the reference, which rejects stateful functors
 * (open_addressing_impl.h:527-532): seed 0, and seed 1 for the second hash
 * of DoubleHashing (probing_scheme.h:99). */
typedef enum mxco_hash {
    MXCO_HASH_MURMUR3_32 = 0, /* murmurhash3_32<Key>, StaticMap default */
    MXCO_HASH_XXHASH_32 = 1,  /* xxhash_32<Key>, StaticSet default */
    MXCO_HASH_XXHASH_64 = 2,  /* xxhash_64<Key> */
    MXCO_HASH_FMIX_32 = 3,    /* murmurhash3_fmix32<Key>, 4-byte keys */
    MXCO_HASH_FMIX_64 = 4     /* murmurhash3_fmix64<Key>, 8-byte keys */
} mxco_hash;

/* KeyEqual: integral keys compare bitwise (== on the integer), floating
 * keys with operator== on float/double (utility/math_utils.h:67).  The
 * sentinel test is always bitwise (detail/bitwise_compare.h:133-145). */
typedef enum mxco_key_kind { MXCO_KEY_INTEGRAL = 0, MXCO_KEY_FLOAT = 1 } mxco_key_kind;

/* Built-in stencil predicates.  The reference takes a Predicate *type*
 * default-constructed on the device (kernels.h:66); through a C ABI the
 * predicate is one of these, applied to each stencil element. */
typedef enum mxco_pred {
    MXCO_PRED_ALWAYS = 0,  /* AlwaysTrue (open_addressing_impl.h:31) */
    MXCO_PRED_NONZERO = 1, /* stencil != 0 */
    MXCO_PRED_ODD = 2,     /* stencil % 2 != 0 */
    MXCO_PRED_EVEN = 3     /* stencil % 2 == 0 */
} mxco_pred;

typedef struct mxco_stencil {
    const void* data;   /* device pointer, one element per input */
    uint32_t elem_size; /* 1, 4 or 8 bytes (signed integers) */
    uint32_t pred;      /* mxco_pred */
} mxco_stencil;

/* ---------------------------------------------------------------------- */
/* Open addressing: StaticMap (value_size > 0) and StaticSet (== 0)       */
/* ---------------------------------------------------------------------- */

typedef struct mxco_oa_desc {
    uint32_t key_size;    /* 4 or 8 */
    uint32_t value_size;  /* 0 (set), 4 or 8; slot = Pair<K,V> layout (pair.h:41) */
    uint32_t key_kind;    /* mxco_key_kind */
    uint32_t bucket_size; /* Storage<B>, 1..64 (storage.h:16-20) */
    uint32_t probing;     /* mxco_probing */
    uint32_t hash;        /* mxco_hash */
    uint32_t group_size;  /* lanes per key: 0 = tuned default, or 1, 2, 4, 8 */
    uint32_t value_kind;  /* mxco_key_kind of the payload (read by DynamicMap's Find merge) */
} mxco_oa_desc;

typedef struct mxco_oa mxco_oa;

/* Capacity the container will use for a requested capacity
 * (MakeValidExtent, detail/extent/extent.inl:21-63). */
MXCO_API mxco_status mxco_valid_extent(uint64_t requested, uint32_t bucket_size, uint32_t probing,
                                       uint64_t* capacity);

/* StaticMap(capacity, emptyKey, emptyValue, ..., stream) static_map.h:78-84,
 * StaticSet(capacity, emptyKey, ..., stream) static_set.h:73-78.
 * empty_value is ignored for sets.  Allocates the slot array and clears it
 * (synchronous, as the reference's constructor, open_addressing_impl.h:77-94). */
MXCO_API mxco_status mxco_oa_create(const mxco_oa_desc* desc, uint64_t capacity, const void* empty_key,
                                    const void* empty_value, hipStream_t stream, mxco_oa** out);
/* ~StaticMap (frees the slot array). */
MXCO_API mxco_status mxco_oa_destroy(mxco_oa* table);
/* Capacity() static_map.h:584, Data() static_map.h:604. */
MXCO_API mxco_status mxco_oa_capacity(const mxco_oa* table, uint64_t* capacity);
MXCO_API mxco_status mxco_oa_data(const mxco_oa* table, void** slots);
MXCO_API mxco_status mxco_oa_slot_size(const mxco_oa* table, uint32_t* slot_size);

/* SetErasedKey(erasedKey, stream) static_map.h:97. */
MXCO_API mxco_status mxco_oa_set_erased_key(mxco_oa* table, const void* erased_key, hipStream_t stream);

/* Clear / ClearAsync static_map.h:318, 331 (open_addressing_impl.h:432-452). */
MXCO_API mxco_status mxco_oa_clear(mxco_oa* table, hipStream_t stream, int sync);

/* Insert / InsertAsync / InsertIf / InsertIfAsync
 * (static_map.h:116,131,216,237; open_addressing_impl.h:102-221).
 * pairs: n slot-layout elements.  stencil == NULL means AlwaysTrue.
 * sync != 0: waits and stores the number of inputs not inserted
 * (duplicates, the reserved empty key, table full) in *num_failed.
 * A NULL pairs/stencil pointer counts every input as failed (sync) or does
 * nothing (async), as the reference does. */
MXCO_API mxco_status mxco_oa_insert(mxco_oa* table, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t stream, int sync, uint64_t* num_failed);

/* InsertOrAssign(values, n, stream[, assignedNum]) / InsertOrAssignAsync
 * (static_map.h:151-175; open_addressing_impl.h:113-157).  Maps only. */
MXCO_API mxco_status mxco_oa_insert_or_assign(mxco_oa* table, const void* pairs, uint64_t n, hipStream_t stream,
                                              int sync, uint64_t* num_failed, uint64_t* num_assigned);

/* AssignIfAsync<StencilT, Predicate> (static_map.h:189; kernels.h:185-210):
 * overwrites the payload of keys already present, never inserts.  Maps only. */
MXCO_API mxco_status mxco_oa_assign_if(mxco_oa* table, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                       hipStream_t stream, int sync);

/* InsertAndFind / InsertAndFindAsync (static_map.h:256, 273;
 * open_addressing_ref_impl.h:302-330): out_found[i] = payload now stored
 * for the key (map) or the stored key (set), out_inserted[i] = 1 when this
 * input inserted it. */
MXCO_API mxco_status mxco_oa_insert_and_find(mxco_oa* table, const void* pairs, void* out_found,
                                             uint8_t* out_inserted, uint64_t n, hipStream_t stream, int sync);

/* Find / FindAsync / FindIf / FindIfAsync (static_map.h:350, 369, 395, 419;
 * kernels.h:314-347): out[i] = payload (map) or stored key (set) of
 * keys[i], or the empty value/key on a miss or a false predicate. */
MXCO_API mxco_status mxco_oa_find(mxco_oa* table, const void* keys, const mxco_stencil* stencil, void* out,
                                  uint64_t n, hipStream_t stream, int sync);

/* Contains / ContainsAsync / ContainsIf / ContainsIfAsync
 * (static_map.h:439, 459, 485, 509; kernels.h:350-377): out[i] = 1 byte bool. */
MXCO_API mxco_status mxco_oa_contains(mxco_oa* table, const void* keys, const mxco_stencil* stencil, uint8_t* out,
                                      uint64_t n, hipStream_t stream, int sync);

/* Count(keys, n, stream) static_map.h:575 (kernels.h:428-467): number of
 * query keys present, synchronous. */
MXCO_API mxco_status mxco_oa_count(mxco_oa* table, const void* keys, uint64_t n, hipStream_t stream,
                                   uint64_t* count);

/* Erase / EraseAsync (static_map.h:292, 307; open_addressing_ref_impl.h:341-379):
 * writes the tombstone set by mxco_oa_set_erased_key, or an empty slot when
 * none is set.  sync != 0 stores the number of keys not erased. */
MXCO_API mxco_status mxco_oa_erase(mxco_oa* table, const void* keys, uint64_t n, hipStream_t stream, int sync,
                                   uint64_t* num_failed);

/* Whether an Insert of n pairs (no stencil) on `stream` takes the region
 * build (no reference counterpart; a diagnostic).  Batches of at least 2^20
 * pairs and half the capacity, on tables of more than 256 MiB of slots
 * (MXCO_OA_REGION_INSERT_MIN_BYTES), under LinearProbing are partitioned by
 * 64 KiB slot region and inserted region by region in LDS, with the CAS
 * walk's results (MXCO_OA_REGION_INSERT=0 turns this off).  The table keeps a
 * workspace of 2 x n slots for it, and concurrent bulk calls on one table,
 * on any streams, need external serialization (as the reference's
 * StaticMultimap states for its calls). */
MXCO_API mxco_status mxco_oa_insert_path(const mxco_oa* table, uint64_t n, hipStream_t stream, int* region_build);

/* Probe statistics of Find's walk over `keys` (no reference counterpart;
 * a diagnostic for roofline accounting): *lines = distinct 128-byte lines
 * and *slots = slots the reference's sequential walk reads
 * (open_addressing_ref_impl.h:568-605), summed over the batch.  Synchronous. */
MXCO_API mxco_status mxco_oa_probe_stats(mxco_oa* table, const void* keys, uint64_t n, hipStream_t stream,
                                         uint64_t* lines, uint64_t* slots);
/* Insert by the CAS walk (never the region build), counting its
 * compare-and-swaps: *cas_made attempted on empty slots and *cas_lost lost to
 * another key, after which the walk goes on at the next slot (no reference
 * counterpart; a diagnostic: the atomic retry rate is lost / made).  The
 * table receives the pairs as Insert does. */
MXCO_API mxco_status mxco_oa_insert_stats(mxco_oa* table, const void* pairs, uint64_t n, hipStream_t stream,
                                          uint64_t* cas_lost, uint64_t* cas_made);

/* Caller-compiled probe kernels: the reference default-constructs any
 * stateless KeyEqual and hasher type on the device
 * (include/detail/open_addressing/kernels.h:64-66).  A caller's types exist
 * only in its own translation unit, so the C++ header instantiates the probe
 * kernels there (include/mxco/detail/caller_probe.h) and registers this
 * launcher on the table: every later probing operation calls it with the
 * operation code, the prepared kernel arguments (struct OaArgs of
 * include/mxco/detail/oa_kernels.h), a grid for kCallerGroup lanes per key,
 * the block size and the stream, and it returns a hipError_t.  Argument
 * checks, counters and synchronization stay in the library; the region
 * build is never taken for such a table.  A null launcher restores the
 * library's kernels. */
typedef int (*mxco_oa_probe_launcher)(int op, const void* kernel_args, uint32_t grid, uint32_t block,
                                      hipStream_t stream);
MXCO_API mxco_status mxco_oa_set_probe_launcher(mxco_oa* table, mxco_oa_probe_launcher launch);

/* ---------------------------------------------------------------------- */
/* Batched tables (no single reference counterpart: the batched form of   */
/* StaticMap::Insert / Find / Contains over many tables, each with the    */
/* reference's per-table results).                                        */
/* ---------------------------------------------------------------------- */

typedef struct mxco_batch mxco_batch;
typedef struct mxco_batch_plan mxco_batch_plan;

/* One work item: `count` inputs of table `table`, starting at element
 * `offset` of the plan's input (and output) arrays. */
typedef struct mxco_batch_item {
    uint32_t table;
    uint32_t reserved;
    uint64_t offset;
    uint64_t count;
} mxco_batch_item;

/* MXCO_BATCH_BUILD: Clear of every table of the batch followed by Insert of
 * the plan's items, as one operation (each table in at most one item).  With
 * LinearProbing the tables are built region by region in LDS; its results
 * (Find / Contains outputs, per-item failures) equal mxco_batch_clear + an
 * insert plan's.  No other operation may touch the batch while it runs. */
enum { MXCO_BATCH_INSERT = 0, MXCO_BATCH_FIND = 1, MXCO_BATCH_CONTAINS = 2, MXCO_BATCH_BUILD = 3 };

/* num_tables tables of one descriptor, table t with requested capacity
 * capacities[t] (rounded as mxco_valid_extent), all in one slot arena. */
MXCO_API mxco_status mxco_batch_create(const mxco_oa_desc* desc, uint32_t num_tables, const uint64_t* capacities,
                                       const void* empty_key, const void* empty_value, hipStream_t stream,
                                       mxco_batch** out);
MXCO_API mxco_status mxco_batch_destroy(mxco_batch* batch);
/* Capacity and slot array of one table (the layout of StaticMap::Data()). */
MXCO_API mxco_status mxco_batch_table(const mxco_batch* batch, uint32_t table, uint64_t* capacity, void** slots);
/* Clears every table with one launch (capture-safe). */
MXCO_API mxco_status mxco_batch_clear(mxco_batch* batch, hipStream_t stream);
/* A device work list for one operation over host items; group_size 0 = default. */
MXCO_API mxco_status mxco_batch_plan_create(mxco_batch* batch, int op, const mxco_batch_item* items,
                                            uint32_t num_items, uint32_t group_size, mxco_batch_plan** out);
MXCO_API mxco_status mxco_batch_plan_destroy(mxco_batch_plan* plan);
/* One persistent-kernel launch of the plan (asynchronous, capture-safe).
 * Insert reads slot-layout pairs; Find writes payloads (or keys for sets);
 * Contains writes 1-byte flags. */
MXCO_API mxco_status mxco_batch_run(mxco_batch_plan* plan, const void* in, void* out, uint8_t* out2,
                                    hipStream_t stream);
/* Per-item failure counts of the last insert run (synchronous). */
MXCO_API mxco_status mxco_batch_failures(mxco_batch_plan* plan, hipStream_t stream, uint64_t* failed_per_item);
/* Autotuner: times the plan under every lane-group width (1/2/4/8) x
 * workgroups per CU x threads per workgroup (64/128/256), and for Find /
 * Contains plans also the wave-staged kernel (256 threads), and keeps the
 * fastest (results do not depend on the choice).  It runs the plan several
 * times; an insert plan's tables are cleared before each timing and are left
 * as one run leaves them, so tune an insert plan before the find plans that
 * read its tables.  A BUILD plan is only timed. */
MXCO_API mxco_status mxco_batch_plan_tune(mxco_batch_plan* plan, const void* in, void* out, uint8_t* out2,
                                          hipStream_t stream, float* best_ms);
MXCO_API mxco_status mxco_batch_plan_config(mxco_batch_plan* plan, uint32_t* group_size, uint32_t* blocks_per_cu);
MXCO_API mxco_status mxco_batch_plan_set_config(mxco_batch_plan* plan, uint32_t group_size, uint32_t blocks_per_cu);
/* Threads per workgroup of the plan's probe kernel (64, 128, 256) and, for
 * Find / Contains, whether it is the wave-staged kernel (which takes 256). */
MXCO_API mxco_status mxco_batch_plan_launch_shape(mxco_batch_plan* plan, uint32_t* block_size, uint32_t* wave_staged);
MXCO_API mxco_status mxco_batch_plan_set_launch_shape(mxco_batch_plan* plan, uint32_t block_size,
                                                      uint32_t wave_staged);

/* ---------------------------------------------------------------------- */
/* StaticMultimap (include/static_multimap.h:88-556): signed int32/int64  */
/* keys and values, duplicates kept.  All calls are synchronous, as in    */
/* the reference (docs/static_multimap_API文档和使用示例.md:23).            */
/* ---------------------------------------------------------------------- */

typedef struct mxco_mm mxco_mm;

/* StaticMultimap(capacity, emptyKey, emptyValue, stream): capacity > 0 pairs
 * (Capacity() returns it unchanged); the table starts empty. */
MXCO_API mxco_status mxco_mm_create(uint32_t key_size, uint32_t value_size, uint64_t capacity, const void* empty_key,
                                    const void* empty_value, hipStream_t stream, mxco_mm** out);
MXCO_API mxco_status mxco_mm_destroy(mxco_mm* map);
/* Capacity() / Size(stream) / Clear(stream) (static_multimap.h:411-428). */
MXCO_API mxco_status mxco_mm_capacity(const mxco_mm* map, uint64_t* capacity);
MXCO_API mxco_status mxco_mm_size(mxco_mm* map, hipStream_t stream, uint64_t* size);
MXCO_API mxco_status mxco_mm_clear(mxco_mm* map, hipStream_t stream);
/* Insert / InsertIf<Stencil, Predicate> (static_multimap.h:323-364, 437-451):
 * checks every selected pair for the reserved key/value first
 * (MXCO_INVALID_ARGUMENT, nothing inserted), then admits the input-order
 * prefix of the selected pairs that fits; *num_failed = selected - admitted. */
MXCO_API mxco_status mxco_mm_insert(mxco_mm* map, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t stream, uint64_t* num_failed);
/* Contains / ContainsIf (static_multimap.h:452-460): 1-byte bool per key. */
MXCO_API mxco_status mxco_mm_contains(mxco_mm* map, const void* keys, const mxco_stencil* stencil, uint8_t* out,
                                      uint64_t n, hipStream_t stream);
/* Find / FindIf (static_multimap.h:461-470): smallest value of the key, or
 * the empty value on a miss or a false predicate. */
MXCO_API mxco_status mxco_mm_find(mxco_mm* map, const void* keys, const mxco_stencil* stencil, void* out_values,
                                  uint64_t n, hipStream_t stream);
/* Count(keys, n, stream) (static_multimap.h:475-495): total matches. */
MXCO_API mxco_status mxco_mm_count(mxco_mm* map, const void* keys, uint64_t n, hipStream_t stream, uint64_t* count);
/* Retrieve(keys, n, probeOut, valueOut, outputCapacity, stream)
 * (static_multimap.h:500-513): every match grouped by query index, values
 * of one query ascending, probe_out holding the query key; MXCO_LENGTH_ERROR
 * (nothing written) when the matches exceed output_capacity. */
MXCO_API mxco_status mxco_mm_retrieve(mxco_mm* map, const void* keys, uint64_t n, void* probe_out, void* value_out,
                                      uint64_t output_capacity, hipStream_t stream, uint64_t* num_out);
/* RetrieveAll(keyOut, valueOut, outputCapacity, stream) (static_multimap.h:523-555):
 * the stored pairs as a multiset (order not part of the contract). */
MXCO_API mxco_status mxco_mm_retrieve_all(mxco_mm* map, void* key_out, void* value_out, uint64_t output_capacity,
                                          hipStream_t stream, uint64_t* num_out);

/* ---------------------------------------------------------------------- */
/* DynamicMap (include/dynamic_map.h:26-625): StaticMap submaps grown on   */
/* demand (room floor(0.6 c) - 10000 per submap, doubling capacities),     */
/* exact deduplication across submaps, erase by tombstone.                 */
/* ---------------------------------------------------------------------- */

typedef struct mxco_dm mxco_dm;

/* InsertMode (dynamic_map.h:36): exact deduplication, or unique keys appended. */
enum { MXCO_DM_EXACT_DEDUP = 0, MXCO_DM_APPEND_UNIQUE = 1 };

/* DynamicMap(initialCapacity, emptyKey, emptyValue, erasedKey, ..., stream)
 * (dynamic_map.h:92-130): maps only; MXCO_INVALID_ARGUMENT when emptyKey ==
 * erasedKey. */
MXCO_API mxco_status mxco_dm_create(const mxco_oa_desc* desc, uint64_t initial_capacity, const void* empty_key,
                                    const void* empty_value, const void* erased_key, hipStream_t stream, mxco_dm** out);
MXCO_API mxco_status mxco_dm_destroy(mxco_dm* map);
/* Size() / Capacity() (sum of submap capacities) / NumSubmaps() (dynamic_map.h:133-137). */
MXCO_API mxco_status mxco_dm_info(const mxco_dm* map, uint64_t* size, uint64_t* capacity, uint64_t* num_submaps);
/* Reserve(n, stream) (dynamic_map.h:142-165). */
MXCO_API mxco_status mxco_dm_reserve(mxco_dm* map, uint64_t n, hipStream_t stream);
/* Insert(values, n, stream[, mode]) (sync != 0, dynamic_map.h:168-219): the
 * number of pairs newly inserted; InsertAsync (sync == 0, :221-235) and
 * MXCO_DM_APPEND_UNIQUE count every pair as inserted.  NULL / 0: nothing. */
MXCO_API mxco_status mxco_dm_insert(mxco_dm* map, const void* pairs, uint64_t n, int mode, int sync,
                                    hipStream_t stream, uint64_t* inserted);
/* InsertOrAssign(values, n, stream) (dynamic_map.h:237-259): new keys inserted (assignments not counted). */
MXCO_API mxco_status mxco_dm_insert_or_assign(mxco_dm* map, const void* pairs, uint64_t n, hipStream_t stream,
                                              uint64_t* inserted);
/* Find / Contains (dynamic_map.h:262-315): every non-empty submap, merged. */
MXCO_API mxco_status mxco_dm_find(mxco_dm* map, const void* keys, void* out, uint64_t n, hipStream_t stream);
MXCO_API mxco_status mxco_dm_contains(mxco_dm* map, const void* keys, uint8_t* out, uint64_t n, hipStream_t stream);
/* Erase (sync != 0: the number erased, dynamic_map.h:318-335) / EraseAsync (:337-347, sizes unchanged). */
MXCO_API mxco_status mxco_dm_erase(mxco_dm* map, const void* keys, uint64_t n, int sync, hipStream_t stream,
                                   uint64_t* erased);
/* Clear(stream) (dynamic_map.h:349-357): every submap cleared, submaps kept. */
MXCO_API mxco_status mxco_dm_clear(mxco_dm* map, hipStream_t stream);
/* Submap i (owned by the DynamicMap), e.g. for its slot array. */
MXCO_API mxco_status mxco_dm_submap(const mxco_dm* map, uint32_t index, mxco_oa** submap);

/* ---------------------------------------------------------------------- */
/* RoaringBitmap (include/roaring_bitmap.h:24-143): membership of uint32 /  */
/* uint64 keys in a bitmap given in the Roaring portable serialization.    */
/* ---------------------------------------------------------------------- */

typedef struct mxco_rb mxco_rb;

/* RoaringBitmap(bitmap[, bitmapBytes], allocator, stream) (roaring_bitmap.h:58-77):
 * key_size 4 (uint32) or 8 (uint64, the 64-bit portable bucket format);
 * bitmap is host memory, bitmap_bytes its length (SIZE_MAX: not known, the
 * caller vouches for complete input).  Malformed, truncated or unordered
 * input: MXCO_INVALID_ARGUMENT. */
MXCO_API mxco_status mxco_rb_create(uint32_t key_size, const void* bitmap, size_t bitmap_bytes, hipStream_t stream,
                                    mxco_rb** out);
MXCO_API mxco_status mxco_rb_destroy(mxco_rb* bitmap);
/* Size(), SizeBytes(), Data() and the device view behind Ref() (roaring_bitmap.h:111-121):
 * bytes is the device copy of the serialized bitmap; the other pointers are
 * the device search image that mxco/detail/roaring.h describes and
 * rb_contains searches (RbBucket table, dense container tables, container
 * keys and descriptors, aligned container data). */
typedef struct mxco_rb_view {
    const uint8_t* bytes;
    const void* buckets;
    const uint64_t* dense;
    const uint16_t* container_keys;
    const uint64_t* container_descs;
    const uint8_t* image;
    uint32_t num_buckets;
    uint32_t reserved;
    uint64_t num_keys;   /* Size() */
    uint64_t size_bytes; /* SizeBytes() */
} mxco_rb_view;
MXCO_API mxco_status mxco_rb_info(const mxco_rb* bitmap, mxco_rb_view* view);
/* Contains / ContainsAsync(keys, outputValues, keyNum, stream) (roaring_bitmap.h:95-108): 1-byte bool per key.
 * keyNum 0 returns at once; null keys or outputValues, or a moved-from
 * (null) bitmap: MXCO_INVALID_ARGUMENT. */
MXCO_API mxco_status mxco_rb_contains(const mxco_rb* bitmap, const void* keys, uint8_t* out, uint64_t n,
                                      hipStream_t stream, int sync);

/* ---------------------------------------------------------------------- */
/* StaticMultiset (include/static_multiset.h:36-575): int32 / int64 keys,  */
/* exact multiplicities, accepted keys kept in insertion order.           */
/* ---------------------------------------------------------------------- */

typedef struct mxco_ms mxco_ms;

/* Query kinds of mxco_ms_query. */
enum {
    MXCO_MS_CONTAINS = 0,        /* Contains / ContainsIf: 1-byte 0/1 */
    MXCO_MS_FIND = 1,            /* Find / FindIf: the key, or the empty key */
    MXCO_MS_COUNT_EACH = 2,      /* CountEach: uint64 multiplicity */
    MXCO_MS_COUNT_EACH_OUTER = 3 /* CountEachOuter: uint64 multiplicity, zero written as one */
};

/* StaticMultiset(capacity, emptyKey, stream) (static_multiset.h:55-76):
 * capacity elements including duplicates (> 0); key_size 4 or 8. */
MXCO_API mxco_status mxco_ms_create(uint32_t key_size, uint64_t capacity, const void* empty_key, hipStream_t stream,
                                    mxco_ms** out);
MXCO_API mxco_status mxco_ms_destroy(mxco_ms* set);
/* Capacity() (static_multiset.h:127-130; 0 for NULL) / Size(stream) (:136-141) / Clear(stream) (:146-157). */
MXCO_API mxco_status mxco_ms_capacity(const mxco_ms* set, uint64_t* capacity);
MXCO_API mxco_status mxco_ms_size(mxco_ms* set, hipStream_t stream, uint64_t* size);
MXCO_API mxco_status mxco_ms_clear(mxco_ms* set, hipStream_t stream);
/* Insert / InsertIf<Stencil, Predicate> (static_multiset.h:163-189): keeps
 * the earliest selected non-sentinel keys that fit; *num_failed = selected
 * keys not kept + selected sentinel keys (unselected inputs are not
 * failures).  NULL keys with n > 0: *num_failed = n, nothing changes. */
MXCO_API mxco_status mxco_ms_insert(mxco_ms* set, const void* keys, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t stream, uint64_t* num_failed);
/* Contains / Find / CountEach / CountEachOuter and the If forms
 * (static_multiset.h:194-260): one output per query; unselected queries
 * read as absent. */
MXCO_API mxco_status mxco_ms_query(mxco_ms* set, int kind, const void* keys, const mxco_stencil* stencil, void* out,
                                   uint64_t n, hipStream_t stream);
/* Count(keys, n, stream) (static_multiset.h:290-294): total multiplicity;
 * MXCO_OVERFLOW_ERROR when n * Size() may exceed 64 bits. */
MXCO_API mxco_status mxco_ms_count(mxco_ms* set, const void* keys, uint64_t n, hipStream_t stream, uint64_t* count);
/* Retrieve(keys, n, probeOutput, matchOutput, outputCapacity, stream)
 * (static_multiset.h:296-321): every match in query order (query key as
 * probe and match); MXCO_LENGTH_ERROR, nothing written, when the matches
 * exceed output_capacity. */
MXCO_API mxco_status mxco_ms_retrieve(mxco_ms* set, const void* keys, uint64_t n, void* probe_out, void* match_out,
                                      uint64_t output_capacity, hipStream_t stream, uint64_t* num_out);
/* RetrieveAll(output, outputCapacity, stream) (static_multiset.h:322-345):
 * the kept keys in accepted insertion order. */
MXCO_API mxco_status mxco_ms_retrieve_all(mxco_ms* set, void* out, uint64_t output_capacity, hipStream_t stream,
                                          uint64_t* num_out);

/* ---------------------------------------------------------------------- */
/* Serving engine (no reference counterpart: a request loop over resident */
/* StaticMap / StaticMultimap tables with continuous batching).  Every    */
/* request's result equals running the table's requests one by one in    */
/* arrival order with the calls above.                                    */
/* ---------------------------------------------------------------------- */

typedef struct mxco_server mxco_server;

typedef struct mxco_server_config {
    uint64_t hbm_reserve_bytes; /* admission keeps at least this much HBM free */
    uint64_t hbm_budget_bytes;  /* 0 = no cap on the tables' total bytes */
    uint64_t max_tick_keys;     /* keys folded into one tick (0 = 2^24); a tick also takes at
                                   most 4096 requests, and one larger request runs alone */
    uint32_t group_size;        /* lanes per key of the batched probes (0 = 4) */
    uint32_t blocks_per_cu;     /* grid of a launched tick (0 = 64) */
    uint32_t tick_mode;         /* 0 = StaticMap ticks run in the resident kernel, fed through a
                                   mailbox in host-coherent memory (4- and 8-byte keys and
                                   payloads); 1 = one insert and one read launch per tick */
    uint32_t resident_blocks_per_cu; /* 512-thread workers per CU of the resident kernel, where they
                                        fit (0 = 1) */
    uint8_t empty_key[8];       /* StaticMap sentinels, little-endian */
    uint8_t empty_value[8];
} mxco_server_config;

enum {
    MXCO_REQ_INSERT = 0,      /* StaticMap::Insert: in = pairs, result = failed count */
    MXCO_REQ_FIND = 1,        /* StaticMap::Find: in = keys, out = values */
    MXCO_REQ_CONTAINS = 2,    /* StaticMap::Contains: out = 1-byte flags */
    MXCO_REQ_MM_INSERT = 3,   /* StaticMultimap::Insert: result = failed count */
    MXCO_REQ_MM_FIND = 4,     /* StaticMultimap::Find: out = smallest values */
    MXCO_REQ_MM_COUNT = 5,    /* StaticMultimap::Count: result = matches */
    MXCO_REQ_MM_RETRIEVE = 6  /* StaticMultimap::Retrieve: out = probe keys, out2 = values, result = matches */
};

typedef struct mxco_request {
    uint32_t table;
    uint32_t op;
    const void* in;
    uint64_t n;
    void* out;
    void* out2;
    uint64_t out_capacity;
} mxco_request;

/* All StaticMap tables of a server share map_desc (key/value sizes, bucket,
 * probing, hash); multimap tables are int64 -> int64. */
MXCO_API mxco_status mxco_server_create(const mxco_server_config* cfg, const mxco_oa_desc* map_desc,
                                        mxco_server** out);
MXCO_API mxco_status mxco_server_destroy(mxco_server* server);
/* Admits a resident table when free HBM (and the budget) allow it,
 * MXCO_OUT_OF_MEMORY otherwise. */
MXCO_API mxco_status mxco_server_add_map(mxco_server* server, uint64_t capacity, uint32_t* table);
MXCO_API mxco_status mxco_server_add_multimap(mxco_server* server, uint64_t capacity, const void* empty_key,
                                              const void* empty_value, uint32_t* table);
/* Refuses further requests to the table, waits for those already submitted,
   then frees the table's HBM. */
MXCO_API mxco_status mxco_server_drop_table(mxco_server* server, uint32_t table);
/* Queues a request and returns at once. */
MXCO_API mxco_status mxco_server_submit(mxco_server* server, const mxco_request* request, uint64_t* ticket);
/* Waits for a ticket: its result and submit-to-completion latency. */
MXCO_API mxco_status mxco_server_wait(mxco_server* server, uint64_t ticket, uint64_t* result, double* latency_us);
/* Waits until every submitted request has completed. */
MXCO_API mxco_status mxco_server_drain(mxco_server* server);
/* Open-loop replay: submits request i at arrival_us[i] after the call starts,
 * then waits for all; fills per-request results and latencies (completion
 * time minus scheduled arrival time, so a late submit counts too).  All
 * arrival times 0 is the closed loop: the latency maximum is the makespan. */
MXCO_API mxco_status mxco_server_replay(mxco_server* server, const mxco_request* requests, const double* arrival_us,
                                        uint32_t n, uint64_t* results, double* latency_us);
/* The same open-loop replay served one request at a time by the plain
 * synchronous calls (the unbatched baseline). */
MXCO_API mxco_status mxco_server_replay_plain(mxco_server* server, const mxco_request* requests,
                                              const double* arrival_us, uint32_t n, uint64_t* results,
                                              double* latency_us);
MXCO_API mxco_status mxco_server_stats(mxco_server* server, uint64_t* ticks, uint64_t* tick_requests);
MXCO_API mxco_status mxco_server_table(mxco_server* server, uint32_t table, mxco_oa** map, mxco_mm** multimap);
/* Whether StaticMap ticks run in the resident kernel, how many launches of
 * it have served them so far (one serves ticks until it is idle for 2 ms,
 * 250 ms old, or stopped between ticks) and its grid. */
MXCO_API mxco_status mxco_server_resident_stats(mxco_server* server, uint32_t* resident, uint64_t* launches,
                                                uint32_t* grid);
/* Diagnostic: waits for outstanding requests, ends the resident kernel and
 * copies the timing trace of the last min(max_ticks, 1024) StaticMap ticks,
 * oldest first, 4 words each: wall-clock times at which the kernel began
 * staging the tick, opened it and published its results, then [insert
 * chunks:32 | read chunks:32]; ticks_per_us converts the times. */
MXCO_API mxco_status mxco_server_resident_trace(mxco_server* server, uint64_t* out, uint32_t max_ticks, uint32_t* n,
                                                double* ticks_per_us);

/* ---------------------------------------------------------------------- */
/* BloomFilter (include/bloom_filter.h:37-133) under                      */
/* BloomFilterPolicy<Key, xxhash_64<Key>, uint32_t, W words, P bits>      */
/* (include/bloom_filter_policy.h:54-211).  Keys are 4 or 8 bytes         */
/* (int32/uint32/float or int64/uint64); the hash covers their bytes.     */
/* ---------------------------------------------------------------------- */

typedef struct mxco_bloom mxco_bloom;

/* The policy's parameters that fix the filter bits
 * (bloom_filter_policy.h:54-98): words_per_block W (a power of two, 1..32),
 * pattern_bits P (a multiple of W, at most 64 and 32 * W; P / W bits per
 * word) and the seed of its xxhash_64<Key> hasher.  The Add / Contains
 * layout parameters of the reference policy change no result and are not
 * carried.  DefaultFilterPolicy<.., W> is {W, W}; ArrowFilterPolicy {8, 8}. */
typedef struct mxco_bloom_policy {
    uint32_t words_per_block;
    uint32_t pattern_bits;
    uint64_t seed;
} mxco_bloom_policy;

/* BloomFilter(numBlocks, policy, allocator, stream)
 * (bloom_filter_lifecycle.inl:15-62): num_blocks blocks of W words, cleared.
 * storage: NULL = the library allocates the words and frees them in
 * destroy; else the caller's device memory of num_blocks * W words (what the
 * C++ header's Allocator returns), which destroy leaves to the caller. */
MXCO_API mxco_status mxco_bloom_create_with_policy(uint64_t num_blocks, uint32_t key_size,
                                                   const mxco_bloom_policy* policy, uint32_t* storage,
                                                   hipStream_t stream, mxco_bloom** out);
/* BloomFilter(numBlocks, policy, stream) under the default policy {8, 8, seed}. */
MXCO_API mxco_status mxco_bloom_create(uint64_t num_blocks, uint32_t key_size, uint64_t seed, hipStream_t stream,
                                       mxco_bloom** out);
MXCO_API mxco_status mxco_bloom_destroy(mxco_bloom* filter);
/* Data() / BlockExtent(): device pointer to the W * num_blocks words. */
MXCO_API mxco_status mxco_bloom_data(const mxco_bloom* filter, uint32_t** words, uint64_t* num_blocks);
/* Clear / ClearAsync. */
MXCO_API mxco_status mxco_bloom_clear(mxco_bloom* filter, hipStream_t stream, int sync);
/* Add / AddAsync / AddIf / AddIfAsync (bloom_filter_add.inl): stencil is a
 * byte per key (non-zero = add) or NULL for all keys. */
MXCO_API mxco_status mxco_bloom_add(mxco_bloom* filter, const void* keys, const uint8_t* stencil, uint64_t n,
                                    hipStream_t stream, int sync);
/* Add that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and the same filter words as mxco_bloom_add; the
 * partitioned pass runs in a fixed scratch array the library holds on each
 * GPU (52 MiB) instead of a grown workspace, so a call takes at most
 * 4,194,304 keys (MXCO_LENGTH_ERROR beyond).  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller. */
MXCO_API mxco_status mxco_bloom_add_noalloc(mxco_bloom* filter, const void* keys, const uint8_t* stencil, uint64_t n,
                                            hipStream_t stream, int sync);
/* Contains / ContainsAsync / ContainsIf / ContainsIfAsync
 * (bloom_filter_query.inl): out[i] = 1 byte bool, 0 where the stencil is 0. */
MXCO_API mxco_status mxco_bloom_contains(const mxco_bloom* filter, const void* keys, const uint8_t* stencil,
                                         uint8_t* out, uint64_t n, hipStream_t stream, int sync);
/* Merge (word-wise OR) / Intersect (word-wise AND) (bloom_filter_merge.inl:23-69). */
MXCO_API mxco_status mxco_bloom_merge(mxco_bloom* dst, const mxco_bloom* src, hipStream_t stream, int sync);
MXCO_API mxco_status mxco_bloom_intersect(mxco_bloom* dst, const mxco_bloom* src, hipStream_t stream, int sync);

/* ---------------------------------------------------------------------- */
/* HyperLogLog (include/hyperloglog.h:24-59): int32 or int64 keys,        */
/* precision p in [13, 18] (2^p one-byte registers).  All calls are       */
/* synchronous, as in the reference.                                      */
/* ---------------------------------------------------------------------- */

typedef struct mxco_hll mxco_hll;

/* CreateWithSketchSizeKB / CreateWithStandardDeviation
 * (hyperloglog.inl:116-140): the precision each factory picks. */
MXCO_API mxco_status mxco_hll_precision_for_sketch_kb(uint32_t sketch_kb, uint32_t* precision);
MXCO_API mxco_status mxco_hll_precision_for_stddev(double deviation, uint32_t* precision);
/* CreateWithPrecision(p) (hyperloglog.inl:124-130); the sketch starts cleared. */
MXCO_API mxco_status mxco_hll_create(uint32_t precision, uint32_t key_size, hipStream_t stream, mxco_hll** out);
MXCO_API mxco_status mxco_hll_destroy(mxco_hll* sketch);
/* Data(): device pointer to the 2^p registers. */
MXCO_API mxco_status mxco_hll_data(const mxco_hll* sketch, const uint8_t** registers, uint32_t* precision);
MXCO_API mxco_status mxco_hll_clear(mxco_hll* sketch, hipStream_t stream);
/* Add(keys, count, stream) (hyperloglog.inl:224-313). */
MXCO_API mxco_status mxco_hll_add(mxco_hll* sketch, const void* keys, uint64_t n, hipStream_t stream);
/* Merge(other, stream) (hyperloglog.inl:315-326): register-wise maximum. */
MXCO_API mxco_status mxco_hll_merge(mxco_hll* dst, const mxco_hll* src, hipStream_t stream);
/* The 65-bin rank histogram Estimate() is computed from (kernels.h:74-92). */
MXCO_API mxco_status mxco_hll_histogram(const mxco_hll* sketch, hipStream_t stream, uint32_t* hist65);
/* Estimate(stream) (hyperloglog.inl:328-345): HLL++ in FP64 on the host. */
MXCO_API mxco_status mxco_hll_estimate(const mxco_hll* sketch, hipStream_t stream, uint64_t* estimate);

/* Number of compute units of the current device (the persistent grids and
 * the tuner size themselves from it). */
MXCO_API mxco_status mxco_device_cu_count(int* cus);

/* Measured ceilings (no reference counterpart; measurement entry points).
 * random_read: n reads of `access` bytes (16, 32, 64, 128 or 256) at hashed,
 * aligned offsets within the largest power-of-two number of `access`-byte
 * units in the first `bytes` of buf, eight in flight per lane group; sink
 * receives one word per lane (CUs * 8 * 256 words).
 * stream_copy: dst = src over `bytes` (a multiple of 16). */
MXCO_API mxco_status mxco_diag_random_read(const void* buf, uint64_t bytes, uint32_t access, uint64_t n,
                                           uint64_t* sink, uint64_t sink_words, hipStream_t stream);
MXCO_API mxco_status mxco_diag_stream_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MXCO_H */
