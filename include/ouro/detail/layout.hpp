// Pool geometry, chunk headers and the device-resident manager state.
//
// The pool keeps the reference's partition (device/MemoryInitialization.cuh:156-187,42-58):
//   [manager state, padded to 8 KiB][max_chunks x 8320 B chunks][reuse rings + class queues][locator]
// so the chunk count, the chunk addresses and every page address are the reference's.
// What changes is the manager state itself: every counter that is hit by atomics sits on its own
// 128-byte line (the reference packs ten queue structs back to back, Ouroboros.cuh:75/171), and
// queue-chunk links are 32-bit chunk indices instead of 64-bit pointers.
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifndef OURO_HD
#if defined(__HIPCC__)
#define OURO_HD __host__ __device__
#else
#define OURO_HD
#endif
#endif

namespace ouro {

// ------------------------------------------------------------------ constants (Parameters.h)
constexpr uint32_t kBaseChunk = 8192;          // CHUNK_SIZE (Parameters.h:18)
constexpr uint32_t kMeta = 128;                // CHUNK_METADATA_SIZE (Parameters.h:19)
constexpr uint32_t kStride = kBaseChunk + kMeta;  // Chunk<8192>::size() (Chunk.cuh:23-26)
constexpr uint32_t kPageBits = 9;              // NUM_BITS_FOR_PAGE (Parameters.h:31)
constexpr uint32_t kPageMask = (1u << kPageBits) - 1u;
constexpr uint32_t kReuseEntries = 65536;      // chunk_queue_size (Parameters.h:10)
constexpr uint32_t kStdEntries = 1u << 23;     // page_queue_size (Parameters.h:11)
constexpr uint32_t kVAEntries = 16384;         // virtual_queue_size (Parameters.h:12)
constexpr uint32_t kQCReserve = 32;            // queue-chunk reserve of a virtualized pool
constexpr uint32_t kQCSpots = kBaseChunk / 4;  // QueueChunk::num_spots_ (QueueChunk.cuh:27)
constexpr uint32_t kOldHold = 10;              // LARGEST_OLD_COUNT_VALUE (Parameters.h:13)
constexpr uint32_t kMaskWords = 8;             // 512 pages / 64-bit words (ChunkAccess.cuh:13 uses 16 x u32)
constexpr uint32_t kStateBytes = 8192;         // Ouroboros::size_() for every chain here (Ouroboros.cuh:260)
// Largest pool in chunks: MemoryIndex keeps 23 chunk bits, and the page indices of chunk 2^23 - 1
// would run into the queue sentinels (0xFFFFFFFC..0xFFFFFFFF), so the top index stays unused. About
// 65 GiB of 8 KiB chunks; initialize() rejects larger pools.
constexpr uint32_t kMaxChunks = (1u << 23) - 1u;

// Queue organisation of one instance (InstanceDefinitions.cuh:1-15).
enum class Kind : uint32_t { PQ = 0, CQ = 1, VAPQ = 2, VACQ = 3, VLPQ = 4, VLCQ = 5 };
OURO_HD constexpr bool is_chunk_kind(Kind k) { return k == Kind::CQ || k == Kind::VACQ || k == Kind::VLCQ; }
OURO_HD constexpr bool is_va(Kind k) { return k == Kind::VAPQ || k == Kind::VACQ; }
OURO_HD constexpr bool is_vl(Kind k) { return k == Kind::VLPQ || k == Kind::VLCQ; }
OURO_HD constexpr bool is_std(Kind k) { return k == Kind::PQ || k == Kind::CQ; }

// MemoryIndex (MemoryIndex.cuh:6-52): 23 chunk bits | 9 page bits.
OURO_HD constexpr uint32_t make_index(uint32_t chunk, uint32_t page) { return (chunk << kPageBits) + page; }
OURO_HD constexpr uint32_t index_chunk(uint32_t idx) { return idx >> kPageBits; }
OURO_HD constexpr uint32_t index_page(uint32_t idx) { return idx & kPageMask; }

OURO_HD constexpr uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }
OURO_HD constexpr uint32_t log2_floor(uint32_t v) { return v <= 1 ? 0 : 1 + log2_floor(v >> 1); }

// Size class, QueueIndex::getQueueIndex (Helper.cuh:42-45): max(ceil(log2(size)) - log2(smallest), 0).
OURO_HD inline uint32_t class_of(uint64_t size, uint32_t smallest_log2) {
  if (size <= 1) return 0;
  uint32_t l = 64 - (uint32_t)__builtin_clzll(size - 1);  // ceil(log2(size))
  return l > smallest_log2 ? l - smallest_log2 : 0;
}

// ------------------------------------------------------------------ semaphore word
// BulkSemaphore word (BulkSemaphore.cuh:5-84), one u64 per class, restated as a debt line:
//   count    (40 bits, bias 2^39): published pages not yet claimed, minus claims not yet covered,
//                                 so it goes negative while carves are in flight;
//   expected (24 bits)           : pages of carves promised but not yet published.
// The reference's third field (reserved, 14 bits) has no counterpart: a claim never waits on the
// word, it waits on the queue itself.
struct Sem {
  static constexpr uint32_t kCountBits = 40;
  static constexpr uint64_t kBias = 1ull << 39;
  static constexpr uint32_t kExpShift = 40;
  static constexpr uint64_t kCountMask = (1ull << kCountBits) - 1ull;
  static constexpr uint64_t kExpMask = (1ull << 24) - 1ull;
  OURO_HD static constexpr uint64_t make(int64_t count, uint64_t expected) {
    return (uint64_t)(count + (int64_t)kBias) | (expected << kExpShift);
  }
  OURO_HD static constexpr int64_t count(uint64_t v) { return (int64_t)(v & kCountMask) - (int64_t)kBias; }
  OURO_HD static constexpr uint64_t expected(uint64_t v) { return (v >> kExpShift) & kExpMask; }
  static constexpr uint64_t kInit = kBias;
};

// Pages the class owes its claimants: claims minus (published + promised) pages.
OURO_HD inline int64_t sem_debt(uint64_t v) { return -(Sem::count(v) + (int64_t)Sem::expected(v)); }

// Chunks a claim of `num` pages must carve, given the word it displaced (BulkSemaphore::wait's
// AllocateChunk decision, BulkSemaphore_impl.cuh:236-263, as a closed form): the claim moves the
// debt from d to d + num and owns the chunk boundaries it crosses, ceil(d1 / ppc) - ceil(d0 / ppc)
// with negative debt counted as zero. Claims see disjoint, consecutive stretches of the line
// (one fetch-add each), promises only lower the debt by whole chunks and frees by single pages,
// so promised-but-unpublished chunks always cover ceil(debt / ppc): no stretch is carved twice
// and none is left uncovered. ppc is a power of two.
OURO_HD inline uint32_t sem_share(uint64_t before, int64_t num, int64_t ppc) {
  const int sh = __builtin_ctzll((unsigned long long)ppc);
  const int64_t d0 = sem_debt(before), d1 = d0 + num;
  const int64_t f0 = d0 > 0 ? (d0 + ppc - 1) >> sh : 0;
  const int64_t f1 = d1 > 0 ? (d1 + ppc - 1) >> sh : 0;
  return (uint32_t)(f1 - f0);
}

// ------------------------------------------------------------------ chunk headers (128 B)
// PageChunk (PageIndexChunk.cuh:5-20): {page_size, number_pages}.
struct PageChunkHdr {
  uint32_t page_size;
  uint32_t number_pages;
};
// ChunkIndexChunk (ChunkIndexChunk.cuh:5-20) + ChunkAccess (ChunkAccess.cuh:26-28). The
// availability mask is eight 64-bit words (one wave-wide ballot per word) instead of 16 x u32.
//
// A chunk carries a generation, set when it is carved for a class: a carve counter (26 bits) and the
// class index (low 4 bits), with the top two bits set. A chunk recycled to the reuse ring and carved
// again gets a new one, so an operation that read a chunk's state before the recycle cannot act on
// its next life: the free count and the generation share one 64-bit word that claims change with a
// 64-bit CAS, and queue_pos carries the generation too. (The top bits keep a generation apart from
// every value the same word of a queue chunk's header holds: chunk indices below 2^23 and markers.)
// queue_pos: bit 63 set = queued at the 32-bit position in bits 0-31 (positions wrap, so the states
// below can never equal one); clear = not queued (0), enqueue in flight (1) or recycled (2) in bits
// 0-31. Bits 32-62 hold the generation's low 31 bits in every state (the reference keeps a 32-bit
// index_t and no generation).
constexpr uint32_t kGenBits = 0xC0000000u;
OURO_HD constexpr uint32_t make_gen(uint32_t carve, uint32_t qi) { return kGenBits | ((carve & 0x3FFFFFFu) << 4) | (qi & 15u); }
OURO_HD constexpr uint32_t gen_class(uint32_t gen) { return gen & 15u; }
OURO_HD constexpr uint64_t qp_state(uint32_t gen, uint32_t s) { return ((uint64_t)(gen & 0x7FFFFFFFu) << 32) | s; }
OURO_HD constexpr uint64_t qp_at(uint32_t gen, uint32_t pos) { return (1ull << 63) | qp_state(gen, pos); }
OURO_HD constexpr uint64_t qp_not(uint32_t gen) { return qp_state(gen, 0); }
OURO_HD constexpr uint64_t qp_pending(uint32_t gen) { return qp_state(gen, 1); }
OURO_HD constexpr uint64_t qp_recycled(uint32_t gen) { return qp_state(gen, 2); }
OURO_HD constexpr bool qp_queued(uint64_t qp) { return (qp >> 63) != 0; }
struct ChunkHdr {
  uint32_t page_size;
  uint32_t size;       // pages on this chunk
  int32_t count;       // free pages
  uint32_t gen;        // generation; {count, gen} form one 64-bit word (count_gen)
  uint64_t queue_pos;  // qp_at(gen, position), qp_not / qp_pending / qp_recycled(gen)
  uint64_t mask[kMaskWords];  // bit set = page free
};
static_assert(sizeof(ChunkHdr) <= kMeta, "chunk header exceeds metadata area");
static_assert(offsetof(ChunkHdr, count) % 8 == 0 && offsetof(ChunkHdr, gen) == offsetof(ChunkHdr, count) + 4,
              "count and gen must form one aligned 64-bit word");
OURO_HD inline uint64_t* count_gen(ChunkHdr* h) { return reinterpret_cast<uint64_t*>(&h->count); }
OURO_HD constexpr uint64_t make_cg(int32_t count, uint32_t gen) { return (uint64_t)(uint32_t)count | ((uint64_t)gen << 32); }
OURO_HD constexpr int32_t cg_count(uint64_t w) { return (int32_t)(uint32_t)w; }
OURO_HD constexpr uint32_t cg_gen(uint64_t w) { return (uint32_t)(w >> 32); }
// QueueChunk (QueueChunk.cuh:37-42): slots live in the 8 KiB data area.
struct QueueChunkHdr {
  uint32_t page_size;      // always 0 (CommonChunk)
  uint32_t count;          // (B << 16) | A: A = enqueued, B = enqueued - dequeued
  uint32_t virtual_start;  // first virtual position held by this chunk
  uint32_t next;           // chunk index of the next queue chunk (linked mode), or marker
  uint32_t chunk_index;
};

// ------------------------------------------------------------------ device-resident state
constexpr uint32_t kRecycleOff = 0xFFFFFFFFu;
// Memory (Ouroboros.cuh:8-29): shared by every instance of a chain.
struct alignas(128) Memory {
  uint8_t* d_memory;
  uint8_t* d_data;
  uint32_t* locator;        // ChunkLocator::d_chunk_flags
  // One byte per base chunk, written when a chunk is carved for a class: log2 of its page size in
  // bits 0-4 and the base chunk's offset inside its run in bits 5-7; 0 = no data chunk. free() reads
  // this dense array (128 chunks per cache line) instead of the locator and the chunk header.
  uint8_t* chunk_log2;
  uint64_t allocation_size;
  uint64_t locator_bytes;
  uint32_t max_chunks;
  uint32_t variant;
  uint32_t initialized;
  uint32_t qc_reserve;      // chunks at the top of the pool that only queue chunks may take
  uint32_t qc_add;          // base chunks per queue chunk: the first instance's chunk run length
  uint32_t stats_on;        // 1: count the statistics below (set by the host)
  uint64_t pool_bytes;      // the whole allocation: the partition above plus the chunk class bytes
  uint64_t instantiation_size;  // initialize()'s arguments, for reinitialize()
  uint32_t add_beg, add_end;
  // chunk-queue recycling (ChunkQueue_impl.cuh:163-199): a chunk whose pages are all free again goes
  // back to the reuse ring when its class queue holds more than recycle_min chunks; kRecycleOff
  // (the default) never recycles, as the reference whose fill level is never reached
  uint32_t recycle_min;
  alignas(128) uint32_t next_free_chunk;
  uint32_t qc_next;         // bump pointer inside the queue-chunk reserve
  uint32_t carves;          // chunk-queue carve counter (generations)
  alignas(128) uint32_t error;
  // Statistics (Statistics.h:5-17): pageAllocCount, pageFreeCount (requests served by / returned to
  // an instance), chunkAllocationCount (chunks handed out, fresh or reused, queue chunks included),
  // chunkReuseCount (chunks returned to a reuse ring)
  alignas(128) unsigned long long stats[4];
};

// IndexQueue (queues/Queue.h:5-24): chunk reuse ring, 65,536 entries per instance.
struct alignas(128) ReuseRing {
  uint32_t* ring;
  uint32_t size;
  int32_t count;
  alignas(128) uint32_t front;
  alignas(128) uint32_t back;
};

// One size class: the per-class queue struct of every queue kind
// (PageQueue.cuh:18-23, ChunkQueue.cuh:20-26, PageQueueVA.cuh:20-26, PageQueueVL.cuh:20-29, ...).
struct alignas(128) ClassQueue {
  uint64_t sem;          // BulkSemaphore
  uint32_t* ring;        // std: element ring; VA: queue-chunk index ring; VL: unused
  uint32_t ring_size;    // power of two
  uint32_t page_size;
  uint32_t ppc;          // pages per chunk
  uint32_t qindex;
  alignas(128) uint32_t front;
  uint32_t front_qc;     // VL front_ptr_
  alignas(128) uint32_t back;
  uint32_t back_qc;      // VL back_ptr_
  alignas(128) int32_t count;   // chunk queues: chunks in the queue (count_)
  uint32_t carved;       // pages (page queues) or chunks (chunk queues) carved into this class
  uint64_t scan;         // chunk queues: page cursor; scan start = queue position scan / ppc
  uint32_t old_qc;       // VL old_ptr_
  uint32_t old_count;    // VL old_count_
};

// ------------------------------------------------------------------ pool layout
// Restates Ouroboros::initialize (MemoryInitialization.cuh:156-166) and initNew (:42-58).
struct PoolLayout {
  uint64_t max_chunks;
  uint64_t allocation_size;
  uint64_t total_required;
  uint64_t locator_bytes;
  uint64_t data_offset;  // d_data - d_memory
  uint64_t queue_offset; // lowest byte of the queue region
};

struct InstanceGeometry {
  uint32_t kind;         // Kind
  uint32_t num_classes;
  uint64_t queue_entries;  // entries per class queue array (0 for VL)
};

OURO_HD inline uint64_t queue_bytes(const InstanceGeometry& g) {
  return align_up(g.queue_entries * 4ull, kBaseChunk);
}
OURO_HD inline uint64_t reuse_bytes() { return align_up((uint64_t)kReuseEntries * 4ull, kBaseChunk); }

OURO_HD inline PoolLayout compute_layout(uint64_t instantiation_size, const InstanceGeometry* g, int n,
                                         uint64_t state_bytes = kStateBytes) {
  PoolLayout L{};
  uint64_t mm = 0;
  for (int i = 0; i < n; ++i) mm += reuse_bytes() + queue_bytes(g[i]) * g[i].num_classes;
  L.total_required = align_up(state_bytes + mm, kStride);
  const uint64_t aligned = align_up(instantiation_size, kStride);
  L.max_chunks = aligned > L.total_required ? (aligned - L.total_required) / kStride : 0;
  const uint64_t adjacency = align_up(L.max_chunks * kStride, 128);
  L.locator_bytes = align_up((L.max_chunks + 31) / 32 * 4, kStride);
  L.allocation_size = align_up(L.total_required + adjacency + L.locator_bytes, kStride);
  L.data_offset = state_bytes;
  L.queue_offset = L.allocation_size - L.locator_bytes - mm;
  return L;
}

}  // namespace ouro
