/*
 * Sequential host simulation of the Ouroboros queue logic -- the oracle.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's CPU leg as the
 * checker. The product (libouro.so, include/ouro) never links or calls it.
 *
 * It replays one wavefront's malloc / free calls with the ordering rules of the device code:
 * lanes requesting the same class form one group led by the lowest lane; chunk carves run one
 * group at a time in lane order; queue positions inside a group follow lane order; queue-chunk
 * allocation and reclamation are issued in lane order. With a single wavefront there is no other
 * concurrency, so a single-wave device trace must equal this replay bit for bit.
 */
#ifndef OURO_SIM_H
#define OURO_SIM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ouro_sim ouro_sim;

/* variant: 0 PQ, 1 CQ, 2 VAPQ, 3 VACQ, 4 VLPQ, 5 VLCQ (MultiOuro* chains). */
ouro_sim* sim_create(int variant, uint64_t pool_bytes);
void sim_destroy(ouro_sim* s);
/* MemoryInitialization.cuh:156-166 */
void sim_layout(int variant, uint64_t pool_bytes, uint64_t* max_chunks, uint64_t* allocation_size);

/* One wavefront of n <= 64 lanes. sizes[i] == 0 marks an inactive lane. out[i] = byte offset of
 * the block from the chunk region, or UINT64_MAX for nullptr. */
void sim_wave_malloc(ouro_sim* s, int n, const uint32_t* sizes, uint64_t* out);
/* offsets[i] == UINT64_MAX marks a lane that frees nullptr (or is inactive). */
void sim_wave_free(ouro_sim* s, int n, const uint64_t* offsets);
/* on = 1: later wave_malloc calls follow the block-wide call of a one-wavefront workgroup
 * (Ouroboros::malloc_tile): n <= 256 requests, request 64 t + l on lane l of sub-tile t; class
 * leaders claim and carve in class order, then the sub-tiles take their pages in order. */
void sim_set_block_order(ouro_sim* s, int on);
/* Fresh pool only: every class queue starts at position `origin` (a multiple of 2048) instead of 0,
 * so a short trace runs its 32-bit positions across the wrap. Returns 0, or -1 if not applicable. */
int sim_set_origin(ouro_sim* s, uint32_t origin);
/* Aggregation width of later chunk-queue wave mallocs: 64 (default; the group claims of widths 16
 * and 64) or 1 (the per-lane claims of the plain path; defined for waves whose lanes request
 * pairwise different classes, lane l being thread l of its workgroup). Returns 0, or -1. */
int sim_set_agg(ouro_sim* s, int w);
/* Chunk recycling for the chunk-queue kinds (the device's flash_chunks): a free that makes its chunk
 * full returns it to the reuse ring while its class queue holds more than min_queued chunks;
 * 0xFFFFFFFF switches it off (the default). Returns 0, or -1 for page-queue kinds. */
int sim_set_recycling(ouro_sim* s, uint32_t min_queued);
/* enqueueInitialChunk on a chunk taken from the pool: its last `available` pages become free pages of
 * class k (counted over the chain). Returns the chunk, or 0xFFFFFFFF when the pool has none. */
uint32_t sim_preload(ouro_sim* s, uint32_t k, uint32_t available);

typedef struct {
  uint32_t next_free_chunk;
  uint32_t error;
  uint32_t num_classes;
  int64_t sem_count[16];
  uint32_t sem_expected[16];
  uint32_t sem_reserved[16];
  uint32_t front[16];
  uint32_t back[16];
  int32_t queue_count[16];
  uint32_t carved[16];
  int32_t reuse_count[2];
  uint32_t reuse_front[2];
  uint32_t reuse_back[2];
} sim_state;
void sim_get_state(const ouro_sim* s, sim_state* out);

/* Chunk-level views for tests: 0 unused, 1 page chunk, 2 chunk-queue chunk, 3 queue chunk. */
int sim_chunk_info(const ouro_sim* s, uint32_t chunk, uint32_t* page_size, int32_t* count, uint64_t* mask8);
/* Chunk queues: the position a chunk is queued at (UINT64_MAX when not queued, UINT64_MAX - 1 while
 * its enqueue is in flight). */
uint64_t sim_chunk_queue_pos(const ouro_sim* s, uint32_t chunk);
/* Slot values of class k's live window [front, back) (0xFFFFFFFD = removed entry); returns its length. */
uint32_t sim_queue_window(const ouro_sim* s, uint32_t k, uint32_t* out, uint32_t max);

/* Class index / page size / pages-per-chunk closed forms (Helper.cuh:42-65) for a request size. */
int sim_class_of(uint32_t size, uint32_t smallest);

/* BulkSemaphore packed-word arithmetic with the 40/24 split used on the device. */
uint64_t sim_sem_make(int64_t count, uint64_t expected);
int64_t sim_sem_count(uint64_t v);
uint64_t sim_sem_expected(uint64_t v);
/* Chunks a claim of num pages carves when the class word holds (count, expected). */
uint32_t sim_sem_share(int64_t count, int64_t expected, int64_t num, int64_t ppc);

/* ChunkLocator::getChunkIndex (ChunkLocator.cuh:29-51) over a caller-supplied bitmap. */
uint32_t sim_locate(const uint32_t* bits, uint32_t nwords, uint32_t idx);

#ifdef __cplusplus
}
#endif
#endif
