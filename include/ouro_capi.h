/*
 * C ABI of the MI355X Ouroboros allocator (libouro.so).
 *
 * The reference ships templates only: its host interface is Ouroboros<...>::initialize /
 * getDeviceMemoryManager (include/device/Ouroboros.cuh:281,297) and its only caller is the harness
 * in src/main.cu. Device code cannot cross a C ABI, so kernels that call malloc/free from their own
 * threads include include/ouro/instances.hpp (the header-only device API). This library wraps the
 * six MultiOuro* instantiations behind plain pointers and sizes for host drivers, tests and
 * benchmarks. Every function returns 0 on success and never exits the process.
 */
#ifndef OURO_CAPI_H
#define OURO_CAPI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ouro_manager* ouro_handle;

/* Instantiations: MultiOuroPQ ... MultiOuroVLCQ (include/InstanceDefinitions.cuh:17-34). */
enum {
  OURO_PQ = 0,   /* page-based, standard array queue        */
  OURO_CQ = 1,   /* chunk-based, standard array queue       */
  OURO_VAPQ = 2, /* page-based, virtualized array queue     */
  OURO_VACQ = 3, /* chunk-based, virtualized array queue    */
  OURO_VLPQ = 4, /* page-based, virtualized linked list     */
  OURO_VLCQ = 5, /* chunk-based, virtualized linked list    */
  /* pool geometries for the batched form's tuner (no reference counterpart): page-based virtualized
   * array queues on one instance, and MultiOuro chains with other queue-array capacities */
  OURO_VAPQ_C8K_Q10 = 6,   /* 8 KiB chunks, 10 classes (16 B .. 8 KiB)            */
  OURO_VAPQ_C8K_Q8 = 7,    /* 8 KiB chunks, 8 classes (64 B .. 8 KiB)             */
  OURO_VAPQ_C16K_Q10 = 8,  /* 16 KiB chunks, 10 classes (32 B .. 16 KiB)          */
  OURO_VAPQ_C16K_Q8 = 9,   /* 16 KiB chunks, 8 classes (128 B .. 16 KiB)          */
  OURO_VAPQ_CAP4K = 10,    /* MultiOuroVAPQ, 4,096 queue chunks per class (2^23 positions) */
  OURO_PQ_CAP1M = 11,      /* MultiOuroPQ, 2^20-entry class rings                 */
  OURO_NUM_VARIANTS = 12
};

/* Geometry of a variant (no pool needed). */
typedef struct {
  uint32_t kind;            /* OURO_PQ ... OURO_VLCQ */
  uint32_t classes;         /* over the whole chain */
  uint32_t smallest_page;   /* bytes */
  uint32_t largest_page;    /* bytes */
  uint32_t chunk_bytes;     /* the first instance's chunk (its largest page) */
  uint32_t queue_entries;   /* entries of each class queue array of the first instance (0: linked) */
  uint64_t page_capacity;   /* pages (page queues) or chunks (chunk queues) one class queue can hold */
} ouro_variant_info;
int ouro_variant_geometry(int variant, ouro_variant_info* out);

/* Status codes. */
enum {
  OURO_OK = 0,
  OURO_E_INVALID = 1, /* bad argument                                  */
  OURO_E_HIP = 2,     /* a HIP runtime call failed                     */
  OURO_E_NOMEM = 3    /* the pool (or a staging buffer) could not be allocated */
};

/* Sticky device error bits reported in ouro_state.error / ouro_trace_result.error_word. */
enum {
  OURO_ERR_OOM = 1u << 0,
  OURO_ERR_QUEUE_FULL = 1u << 1,
  OURO_ERR_VSTART = 1u << 2,
  OURO_ERR_TIMEOUT = 1u << 3,
  OURO_ERR_BAD_FREE = 1u << 4,
  OURO_ERR_TOO_LARGE = 1u << 5,
  OURO_ERR_REUSE_FULL = 1u << 6,
  OURO_ERR_LOCATOR = 1u << 7,
  OURO_ERR_CORRUPT = 1u << 8
};

#define OURO_MAX_CLASSES 16

/* Snapshot of the device-resident manager state (printFreeResources, device/MemoryQueries.cuh:8-61,
 * as data rather than printf). Classes of all instances are listed in chain order. */
typedef struct {
  uint64_t allocation_size;
  uint64_t max_chunks;
  uint64_t data_offset;
  uint32_t next_free_chunk;
  uint32_t error;
  uint32_t variant;
  uint32_t num_classes;
  uint32_t page_size[OURO_MAX_CLASSES];
  uint32_t pages_per_chunk[OURO_MAX_CLASSES];
  int64_t sem_count[OURO_MAX_CLASSES];
  uint32_t sem_expected[OURO_MAX_CLASSES];
  uint32_t sem_reserved[OURO_MAX_CLASSES];
  uint32_t front[OURO_MAX_CLASSES];
  uint32_t back[OURO_MAX_CLASSES];
  int32_t queue_count[OURO_MAX_CLASSES];
  uint32_t carved[OURO_MAX_CLASSES];
  uint32_t front_qc[OURO_MAX_CLASSES];   /* linked-list queues: first queue chunk */
  uint32_t kind[OURO_MAX_CLASSES];       /* OURO_PQ ... OURO_VLCQ of the class's instance */
  int32_t reuse_count[2];
} ouro_state;

/* Quiescent-state invariant check: after every block has been freed, each class's semaphore,
 * queue and chunk bitmasks must describe exactly the carved pages, all free. */
typedef struct {
  uint64_t violations;      /* total number of violated invariants      */
  uint64_t chunks_checked;
  uint64_t pages_free;      /* pages found free in carved chunks        */
  int32_t first_class;      /* class of the first violation, or -1      */
  int32_t first_chunk;      /* chunk of the first violation, or -1      */
  char message[160];        /* description of the first violation       */
} ouro_check_result;

/* One alloc / write / verify / free round per "round", over `threads` device threads
 * (src/main.cu:146-167, with per-thread sizes). */
typedef struct {
  uint32_t threads;
  uint32_t rounds;
  uint32_t agg;                /* aggregation width: 1 (reference), 16 or 64 */
  uint32_t block;              /* threads per block (0 = 256)                  */
  uint32_t check_rounds;       /* 1 = host invariant check of the live set every round */
  uint32_t latency;            /* 1 = record malloc latency on lane 0 of every wave */
  uint32_t hold_permille;      /* blocks whose free is deferred by one round (per mille) */
  uint32_t pad;
  uint64_t hold_seed;
  const uint32_t* sizes;       /* host [threads] sizes, fixed across rounds ...          */
  const uint32_t* round_sizes; /* ... or host [rounds * threads], overrides `sizes`     */
  uint64_t* offsets_out;       /* optional host [rounds * threads]: byte offset of each
                                  block from the data region, UINT64_MAX for nullptr     */
} ouro_trace_config;

typedef struct {
  double alloc_ms;             /* summed over rounds, hipEvent-timed malloc kernels */
  double free_ms;              /* summed over rounds, hipEvent-timed free kernels   */
  double write_ms;
  double verify_ms;
  uint64_t mallocs;
  uint64_t frees;
  uint64_t null_returns;
  uint64_t verify_errors;      /* words that did not read back the writer's id        */
  uint64_t bounds_errors;      /* blocks outside the data region                      */
  uint64_t overlap_errors;     /* live blocks that overlap (check_rounds only)         */
  uint64_t align_errors;       /* blocks not on a page boundary of their class         */
  uint64_t touch_bytes;        /* algorithmic bytes of the write + verify passes        */
  uint64_t peak_chunks;        /* highest next_free_chunk seen                        */
  uint64_t peak_live_chunks;   /* most chunks holding a live block at a round end (check_rounds) */
  uint64_t ideal_chunks;       /* sum over classes of ceil(live / pages per chunk) at that round */
  uint64_t lat_samples;
  double lat_p50_ns;
  double lat_p99_ns;
  double lat_max_ns;
  uint32_t error_word;
  uint32_t pad;
} ouro_trace_result;

/* ---- lifetime ------------------------------------------------------------------------------ */
/* Ouroboros::initialize(pool_bytes) (device/MemoryInitialization.cuh:123-219) on the current device. */
int ouro_create(int variant, uint64_t pool_bytes, ouro_handle* out);
/* reinitialize(overallocation_factor) (declared, never defined, in device/Ouroboros.cuh:83,179,283):
 * replaces the pool with a fresh one of overallocation_factor times the size given to ouro_create
 * (grow with a factor above 1). Every block handed out before becomes invalid and the device
 * manager moves (call ouro_device_manager again); call it with no block live. If the new pool
 * cannot be allocated the handle has no pool left (OURO_E_NOMEM) and can only be destroyed. */
int ouro_reinitialize(ouro_handle h, double overallocation_factor);
/* Memory::~Memory (device/Ouroboros.cuh:11-15): frees the pool. */
int ouro_destroy(ouro_handle h);
/* getDeviceMemoryManager() (device/Ouroboros.cuh:297): pass to kernels built on include/ouro. */
void* ouro_device_manager(ouro_handle h);
/* Start of the chunk region (Memory::d_data, device/Ouroboros.cuh:18). */
void* ouro_data_base(ouro_handle h);
/* Pool arithmetic without a device (device/MemoryInitialization.cuh:156-166). */
int ouro_layout(int variant, uint64_t pool_bytes, uint64_t* max_chunks, uint64_t* allocation_size);

/* ---- batched device calls ------------------------------------------------------------------- */
/* d_testAllocation (src/main.cu:10-18): d_out[i] = malloc(d_sizes[i]) on n device threads. */
int ouro_malloc_batch(ouro_handle h, const uint32_t* d_sizes, void** d_out, uint32_t n, int agg, void* stream);
/* d_testFree (src/main.cu:55-63): free(d_ptrs[i]) on n device threads (nullptr entries are skipped). */
int ouro_free_batch(ouro_handle h, void* const* d_ptrs, uint32_t n, int agg, void* stream);

/* ---- state and checks ----------------------------------------------------------------------- */
/* printFreeResources (device/MemoryQueries.cuh:8-61) as data. */
int ouro_query(ouro_handle h, ouro_state* out);
/* Invariant checker of a quiescent pool (no reference counterpart): queues, masks and counts back at
 * what the carved chunks account for, no error bit set. */
int ouro_check_quiescent(ouro_handle h, ouro_check_result* out);
/* enqueueInitialChunk (Ouroboros.cuh:291-295; PageQueue_impl.cuh:40-53, ChunkQueue_impl.cuh:72-78)
 * for bulk-loaded consumers: takes one chunk for size class `class_index` (counted over the whole
 * chain) from the pool and makes its last `available_pages` pages free pages of the class; the first
 * pages_per_chunk - available_pages pages belong to the caller, who may hand them back with free().
 * *chunk_out = the chunk index (page p is at data + chunk * 8320 + 128 + p * page size).
 * OURO_E_NOMEM when the pool has no chunk left. Call before any kernel uses the pool concurrently. */
int ouro_preload(ouro_handle h, uint32_t class_index, uint32_t available_pages, uint32_t* chunk_out);

/* Statistics counters (Statistics.h:5-17, incremented as Ouroboros_impl.cuh:20-21,33-34,53-54,66-67
 * and Ouroboros.cuh:113-114 do), compiled in unless the library is built with OURO_STATISTICS=0 and
 * counted only while switched on, one atomic per wave or workgroup and call:
 * out4[0] pageAllocCount   requests an instance served or refused (not the device-heap ones),
 * out4[1] pageFreeCount    blocks returned to an instance,
 * out4[2] chunkAllocationCount  chunks handed out for data or queue chunks, fresh or reused,
 * out4[3] chunkReuseCount  chunks returned to a reuse ring.
 * ouro_set_statistics(h, 1) zeroes the counters and starts counting, (h, 0) stops; both wait for the
 * device. OURO_E_INVALID when the library has no statistics. */
int ouro_set_statistics(ouro_handle h, int on);
/* Chunk recycling (the flash of ChunkQueue::freePage, device/queues/chunk/ChunkQueue_impl.cuh:163-199)
 * for the chunk-queue kinds (CQ, VACQ, VLCQ): a free that makes every page of its chunk free again
 * returns the chunk to the instance's reuse ring, where any class can carve it, whenever its class
 * queue holds more than min_queued chunks (the reference's lower_fill_level, ChunkQueue.cuh:30).
 * OURO_RECYCLE_OFF (the default after ouro_create) switches it off: the reference's own fill level
 * (838,860 queued chunks) is never reached, and its VA and VL chunk queues never recycle
 * (ChunkQueueVA_impl.cuh:229, ChunkQueueVL_impl.cuh:160-192). Waits for the device; OURO_E_INVALID
 * for page-queue kinds, whose chunks are never recycled. */
#define OURO_RECYCLE_OFF 0xFFFFFFFFu
int ouro_set_recycling(ouro_handle h, uint32_t min_queued);
int ouro_statistics(ouro_handle h, uint64_t* out4);

/* Bytes of a snapshot: the reference's pool partition (allocation_size) plus the chunk class bytes
 * and the linked queues' queue-chunk directories that follow it in the same allocation. */
int ouro_snapshot_size(ouro_handle h, uint64_t* bytes);
/* Copy the complete allocator state (the whole allocation, ouro_snapshot_size bytes) to dst, device
 * or host memory. Take it while no kernel uses the pool. (Ouroboros keeps all state in the pool and
 * mirrors the manager struct into it, Ouroboros_impl.cuh:137-155.) */
int ouro_snapshot(ouro_handle h, void* dst);
/* Replace the state of h with a snapshot taken from a pool of the same variant and pool size (device
 * or host memory): the allocator continues exactly where the snapshot was taken. The addresses
 * inside the manager state are rebased onto h's pool. OURO_E_INVALID for a snapshot of another
 * variant or size. No reference counterpart (the reference's updateMemoryManagerDevice,
 * Ouroboros_impl.cuh:148-155, copies only the manager struct of the same pool). */
int ouro_restore(ouro_handle h, const void* src);

/* ---- harness -------------------------------------------------------------------------------- */
/* The harness loop of src/main.cu:146-172: alloc / write / verify / free rounds, kernels timed apart. */
int ouro_run_trace(ouro_handle h, const ouro_trace_config* cfg, ouro_trace_result* out);
/* ---- batched form (no reference counterpart) ------------------------------------------------- */
/* A flat list of allocation requests served by persistent kernels with block-wide aggregation
 * (agg 0: every size class of a workgroup's tile claims its pages with one semaphore operation and
 * one front_ fetch-add) or by the per-thread kernels (agg 1, 16, 64). One round = alloc, write the
 * requester's id into every word, read it back, free; with graph = 1 the round is captured once in
 * a hipGraph and replayed. */
typedef struct {
  uint32_t requests;       /* allocation requests */
  uint32_t rounds;         /* timed rounds */
  uint32_t warmup;         /* untimed rounds before them */
  int32_t agg;             /* 0 = block-wide (persistent kernels), 1 / 16 / 64 = per-wave widths */
  uint32_t block;          /* threads per workgroup: a multiple of 64, at most 1024 */
  uint32_t grid;           /* persistent workgroups (agg 0), 0 = every CU at its occupancy limit */
  int32_t graph;           /* 1 = replay one captured round graph */
  int32_t check;           /* 1 = a validation round first (overlap / alignment of the live set) */
  const uint32_t* sizes;   /* host [requests] */
  uint64_t* offsets_out;   /* optional host [requests]: the validation round's blocks as byte offsets
                              from the data region, UINT64_MAX for nullptr */
  int32_t latency;         /* 1 = one more (untimed) round recording the wall-clock time of each
                              allocation call on lane 0 of every wave */
  uint32_t tile;           /* agg 0: requests per thread and block-wide call, 1 (0), 2 or 4; a tile
                              of tile * block requests is served as one block-wide group per class */
  int32_t kernel_timing;   /* agg 0: 1 = after the timed rounds, up to 20 more rounds of plain launches
                              whose persistent kernels are timed by their own dispatch's events */
  uint32_t pad;
} ouro_batch_config;

typedef struct {
  double alloc_ms;         /* summed over timed rounds, hipEvent-timed alloc kernel */
  double free_ms;          /* summed over timed rounds, hipEvent-timed free kernel */
  double touch_ms;         /* write + verify passes */
  double round_ms;         /* whole rounds (alloc .. free) */
  uint64_t mallocs;        /* requests x timed rounds */
  uint64_t null_returns;   /* over the timed rounds */
  uint64_t verify_errors;
  uint64_t bounds_errors;
  uint64_t overlap_errors; /* validation round */
  uint64_t align_errors;   /* validation round */
  uint64_t touch_bytes;    /* algorithmic bytes of one round's write + verify passes */
  uint32_t error_word;
  uint32_t grid;           /* workgroups of the alloc / free kernels */
  uint32_t peak_chunks;
  uint32_t graph_used;     /* 1 if the rounds were graph replays */
  double lat_p50_ns, lat_p99_ns, lat_max_ns; /* latency round (cfg->latency) */
  uint64_t lat_samples;
  /* cfg->kernel_timing: kernel-only durations of the persistent alloc / free kernels, summed over
     kernel_rounds extra rounds of plain launches after the timed ones, each kernel timed by events
     its own dispatch records (hipExtLaunchKernel); the timed rounds above bracket each kernel, or
     graph node, with separate event packets */
  double alloc_kernel_ms, free_kernel_ms;
  uint64_t kernel_rounds;
} ouro_batch_result;

int ouro_run_batch(ouro_handle h, const ouro_batch_config* cfg, ouro_batch_result* out);

/* ouro_run_batch without any allocation (no reference counterpart): for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same rounds as ouro_run_batch -- the observed round (cfg->check; its blocks in
 * cfg->offsets_out), cfg->warmup rounds, then cfg->rounds rounds whose verify / bounds / null counts
 * are reported -- on the caller's stream (NULL: the null stream), with the request sizes, block
 * pointers and error counters staged in a static device scratch of 12 MiB per GPU instead of
 * buffers the call allocates. At most 2^20 requests (OURO_E_INVALID beyond); graph, latency and
 * kernel_timing must be 0; no timings and no host overlap / alignment check (those fields stay 0).
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU race. */
int ouro_run_batch_static(ouro_handle h, const ouro_batch_config* cfg, void* stream, ouro_batch_result* out);

/* Host trace generator (SplitMix64): dist 0 = fixed size `param`, 1 = uniform over the class
 * sizes 16 << q (q < classes), 2 = Zipf(param) over class ranks. */
int ouro_gen_sizes(uint64_t seed, uint32_t n, int dist, double param, uint32_t classes, uint32_t* out);

/* Batch workload generator (SplitMix64 keyed by seed and workload): workload k has
 * R_k = clamp(round(lognormal(mu, sigma)), lo, hi) requests, a Zipf exponent s_k ~ U[s_lo, s_hi] over
 * the class ranks (P(q) proportional to (q + 1)^-s_k) and 1 .. max_allocs allocations per request.
 * Writes the flat allocation sizes (16 << q), request after request, to sizes[0 .. cap) and the total
 * in *n_out (call with sizes = NULL to size the buffer); requests_out / zipf_out are optional
 * [workloads] arrays. */
typedef struct {
  uint64_t seed;
  uint32_t workloads;
  uint32_t lo, hi;          /* request-count clamp */
  uint32_t max_allocs;      /* allocations per request: 1 .. max_allocs */
  uint32_t classes;
  uint32_t pad;
  double mu, sigma;         /* lognormal request count */
  double s_lo, s_hi;        /* per-workload Zipf exponent range */
} ouro_batch_workload;
int ouro_gen_batch(const ouro_batch_workload* w, uint32_t* sizes, uint64_t cap, uint64_t* n_out,
                   uint32_t* requests_out, double* zipf_out);

/* ---- serving form (no reference counterpart: the reference has no request loop) ---------------
 * A host request loop in front of one device-resident pool. Request r arrives at a Poisson time
 * (rate per second), runs one kernel of 2^e threads (e ~ U[min_log2, max_log2]); thread i
 * allocates a block of class q ~ Zipf(zipf) over the first `classes` class ranks (16 B << q) through
 * the per-thread device malloc of width `agg` or the block-wide malloc (agg 0), fills it with a
 * pattern keyed by (request, i), reads it back and frees it. Continuous batching: whenever one of
 * `slots` launch slots (one stream each) is free, up to `max_batch` requests that have arrived are
 * served by one launch, each workgroup working on a stretch of one request, and launches overlap.
 * slots = 1, max_batch = 1, agg = 1 is the plain per-request launch path. With admit = 1 a request
 * joins a launch only when the pool's free pages, copied asynchronously from the device state, cover
 * the demand of everything in flight plus its own. Latency is arrival to the completion of the
 * request's launch, on the device clock. */
typedef struct {
  double rate;             /* arrivals per second */
  double zipf;             /* class-rank exponent */
  uint64_t seed;
  uint32_t requests;
  uint32_t slots;          /* requests in flight at most (1 .. 64) */
  int32_t agg;             /* 0 = block-wide (malloc_block / free_block), 1 / 16 / 64; 256-thread workgroups */
  uint32_t min_log2, max_log2; /* threads per request: 2^U[min, max], max <= 24 */
  uint32_t classes;        /* size classes drawn (1 .. 10) */
  int32_t admit;           /* 1 = admission by the device free-page count */
  uint32_t max_batch;      /* requests one launch may carry (0 or 1: one request per launch) */
  unsigned long long* results_out; /* optional host [requests][4]: words intact, words not, nulls, threads */
  double* latency_out;     /* optional host [requests] ms */
  uint32_t tile;           /* agg 0: request threads per GPU thread, 1 (0), 2 or 4 */
  uint32_t pad;
} ouro_serve_config;

typedef struct {
  double p50_ms, p99_ms, mean_ms, max_ms;
  double wall_ms;          /* first arrival to last completion */
  double offered_ms;       /* first to last arrival */
  double mops;             /* (mallocs + frees) / wall */
  double achieved_rate;    /* requests per second */
  uint64_t mallocs;
  uint64_t mismatches;     /* words read back wrong, all requests */
  uint64_t nulls;
  uint32_t completed;
  uint32_t max_in_flight;
  uint32_t deferred;       /* requests the admission test held back at least once */
  uint32_t error_word;
  uint32_t launches;       /* kernel launches that served the requests */
  uint32_t pad;
} ouro_serve_result;

int ouro_serve(ouro_handle h, const ouro_serve_config* cfg, ouro_serve_result* out);
/* Threads and intact words request r of `cfg` returns when every allocation succeeds. */
int ouro_serve_expected(const ouro_serve_config* cfg, uint32_t r, uint64_t* threads, uint64_t* words);

/* ---- test hook (no reference counterpart) --------------------------------------------------- */
/* The device semaphore's carve decision (BulkSemaphore_impl.cuh:236-263 as changed here) evaluated
 * on the host: *k = chunks a claim of `num` pages must carve, given the word its fetch-add displaced
 * (`before`); ppc a power of two. sem_make packs (count, expected) the way the device does. */
int ouro_sem_share(uint64_t before, int64_t num, int64_t ppc, uint32_t* k);
uint64_t ouro_sem_make(int64_t count, uint64_t expected);

/* Position origin (test hook, no reference counterpart): on a pool no call has used yet, start every
 * class queue at position `origin` (a multiple of 2048) instead of 0: front_, back_, the chunk-queue
 * page cursor and the initial VA / VL queue chunk. Queue positions are 32-bit and wrap; a trace run
 * from an origin just below 2^32 crosses the wrap in a few rounds and must hand out exactly the
 * blocks it hands out from origin 0. */
int ouro_set_origin(ouro_handle h, uint32_t origin);

/* Layout check for host bindings (no reference counterpart): out[i] = sizeof of the i-th public
 * struct, in the order ouro_variant_info, ouro_state, ouro_check_result, ouro_trace_config,
 * ouro_trace_result, ouro_batch_config, ouro_batch_result, ouro_batch_workload, ouro_serve_config,
 * ouro_serve_result; n <= 10 entries are written. Host only. */
int ouro_abi_sizes(uint32_t* out, uint32_t n);

/* ---- measurement helpers (no reference counterpart) ---------------------------------------- */
/* Peak device-scope atomic throughput: `threads` threads each issue `iters` u32 fetch-adds with
 * return. pattern 0 = every lane its own address (distinct cache lines), 1 = one address per wave,
 * 2 = one address for the whole grid. Returns atomics per second in *ops_per_s. */
int ouro_bench_atomics(uint32_t threads, uint32_t iters, int pattern, double* ops_per_s);
/* Plain device copy of `bytes` bytes (16-byte vector loads/stores), `reps` times; returns GB/s.
 * Used to calibrate the FETCH_SIZE / WRITE_SIZE counters against known traffic. */
int ouro_bench_copy(uint64_t bytes, uint32_t reps, double* gb_per_s);
/* The floor under the batched malloc: a persistent kernel at the launch shape of the allocator's
 * (block, tile, and the grid ouro_run_batch would use, or `grid` if smaller) that loads the n sizes,
 * ranks them per class in LDS with the same barriers, and writes a made-up pointer per request,
 * touching no allocator state. Average kernel time over `rounds` launches (after one warm-up),
 * each timed by its own dispatch's events, in *kernel_ms; the grid used in *grid_out (may be NULL). */
int ouro_bench_floor(ouro_handle h, const uint32_t* sizes, uint32_t n, uint32_t block, uint32_t grid, uint32_t tile,
                     uint32_t rounds, double* kernel_ms, uint32_t* grid_out);

#ifdef __cplusplus
}
#endif
#endif /* OURO_CAPI_H */
