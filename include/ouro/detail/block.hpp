// Workgroup-wide allocation: every thread of a block calls malloc_block / free_block together, and
// requests of one size class are served as one group of up to blockDim.x lanes. The semaphore claim
// and the queue front / back fetch-adds (the same-address atomics of a class) are issued once per
// class per block instead of once per wave group; the per-request work (queue slots, page bits)
// stays per lane. This is the aggregation the batched form uses; the reference aggregates at most a
// warp (BulkSemaphore_impl.cuh:207-299, Utility.cuh:76-95).
//
// Order inside a block is fixed: requests of a class are ranked by wave, then by lane, and class
// leaders run in wave 0 in class order, so a single block replays deterministically.
#pragma once
#include "instance.hpp"

namespace ouro {

constexpr uint32_t kBlockMaxWaves = 16;    // 1024-thread blocks
constexpr uint32_t kBlockMaxClasses = 16;  // size classes over the whole instance chain

// LDS scratch of one workgroup. The tuple list (chunk queues only) holds the claimed page runs.
template <bool kChunks>
struct BlockScratch {
  struct Run {
    uint32_t bc;    // chunk | mask word << 29
    uint32_t r0;    // first rank served by the run
    uint64_t bits;  // claimed pages of that mask word
  };
  uint32_t off[kBlockMaxWaves][kBlockMaxClasses];  // requests of (wave, class), then their rank offset
  uint32_t total[kBlockMaxClasses];
  uint32_t base[kBlockMaxClasses];    // page queues: first queue position; chunk queues: first run
  uint32_t nrun[kBlockMaxClasses];    // chunk queues: runs written
  uint32_t served[kBlockMaxClasses];  // chunk queues: ranks served (the rest get nullptr)
  Run run[kChunks ? 1024 : 1];
};

namespace detail {

// Rank of this lane among the block's requests of class gc (kMarker: no request), after the class
// totals and per-wave offsets are in sm. Ends with the block synchronised.
template <class S>
__device__ inline uint32_t block_rank(S& sm, uint32_t gc, uint32_t nclasses) {
  const uint32_t wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63u) >> 6;
  if (lane() < kBlockMaxClasses) sm.off[wave][lane()] = 0;
  const bool valid = gc != kMarker;
  const uint64_t grp = match<64>(valid ? (uint64_t)gc : ~0ull);
  const uint32_t r = rank_in(grp);
  if (valid && (int)lane() == first(grp)) sm.off[wave][gc] = popc(grp);
  __syncthreads();
  if (wave == 0 && lane() < nclasses) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < nwaves; ++w) {
      const uint32_t c = sm.off[w][lane()];
      sm.off[w][lane()] = run;
      run += c;
    }
    sm.total[lane()] = run;
  }
  __syncthreads();
  return valid ? sm.off[wave][gc] + r : 0u;
}

// ---- page queues
// Wave 0: lane gc (off <= gc < off + NQ) claims its class's block total, the carves run in lane
// order, then one front_ fetch-add per class.
template <class I, class S>
__device__ inline void block_claim_pages(Memory& m, I& inst, ReuseRing& base, uint32_t off, S& sm) {
  const uint32_t l = lane();
  const bool mine = l >= off && l < off + I::kNQ;
  const uint32_t total = mine ? sm.total[l] : 0u;
  const bool lead = total > 0;
  if (!ballot(lead)) return;
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[mine ? l - off : 0u];
  auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_pages(x, lq, k); };
  sem_claim_n(x, cq, lead, total, carve);
  if (lead) sm.base[l] = add(&cq->front, total);
}

template <class I>
__device__ inline void* block_take_page(Memory& m, I& inst, ReuseRing& base, bool here, uint32_t qi, uint32_t pos) {
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[here ? qi : 0u];
  const uint32_t idx = take_lanes<64>(x, cq, here, pos);
  if (!here || idx == kMarker || idx == kNullPage) return nullptr;
  if (!chunk_ok(m, index_chunk(idx)) || index_page(idx) >= ppc_of(x, cq)) {
    set_error(m, ERR_CORRUPT);
    return nullptr;
  }
  return chunk_addr(m, index_chunk(idx)) + kMeta + (uint64_t)index_page(idx) * page_size_of(x, cq);
}

// Wave 0: back_ fetch-add per class of the block's frees.
template <class I, class S>
__device__ inline void block_back(I& inst, uint32_t off, S& sm) {
  const uint32_t l = lane();
  const bool mine = l >= off && l < off + I::kNQ;
  const uint32_t total = mine ? sm.total[l] : 0u;
  if (total) sm.base[l] = add(&inst.q[l - off].back, total);
}

// Wave 0, after every put of the block is done: one signal per class (publish, then signal).
template <class I, class S>
__device__ inline void block_signal(I& inst, uint32_t off, S& sm) {
  const uint32_t l = lane();
  const bool mine = l >= off && l < off + I::kNQ;
  const uint32_t total = mine ? sm.total[l] : 0u;
  if (total) sem_signal(&inst.q[l - off], total);
}

// ---- chunk queues
// Wave 0: lane gc leads a ChunkClaim for its class's block total; every run of pages it claims is
// appended to the class's stretch of sm.run (a stretch of `total` entries starting at the class's
// prefix of totals, enough since every run serves at least one rank).
template <class I, class S>
__device__ inline void block_claim_chunks(Memory& m, I& inst, ReuseRing& base, uint32_t off, S& sm,
                                          uint32_t nclasses) {
  const uint32_t l = lane();
  const bool mine = l >= off && l < off + I::kNQ;
  const uint32_t total = mine ? sm.total[l] : 0u;
  const bool lead = total > 0;
  if (!ballot(lead)) return;
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[mine ? l - off : 0u];
  auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_chunks(x, lq, k); };
  sem_claim_n(x, cq, lead, total, carve);
  // stretch start: prefix of the block totals over the classes before l
  uint32_t start = 0;
  for (uint32_t c = 0; c < nclasses; ++c) start += c < l ? sm.total[c] : 0u;
  ChunkClaim<I> cl(x, cq, lead, total);
  uint32_t nrun = 0;
  while (ballot(lead && !cl.finished())) {
    if (lead && !cl.finished()) {
      uint32_t bc, r0;
      uint64_t bits;
      cl.step(x, cq, bc, bits, r0);
      if (bc != kMarker) {
        sm.run[start + nrun] = {bc, r0, bits};
        ++nrun;
      }
    }
  }
  if (lead) {
    sm.base[l] = start;
    sm.nrun[l] = nrun;
    sm.served[l] = cl.assigned;
  }
  finish_reclaim(x, cq, lead, cl.reclaim);
}

// Page of rank r from the class's runs (binary search on r0).
template <class I, class S>
__device__ inline void* block_page_of(Memory& m, I& inst, const S& sm, bool here, uint32_t gc, uint32_t qi, uint32_t r) {
  if (!here || r >= sm.served[gc]) return nullptr;
  uint32_t lo = sm.base[gc], hi = lo + sm.nrun[gc];  // r0 ascending; find the last run with r0 <= r
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sm.run[mid].r0 <= r) lo = mid;
    else hi = mid;
  }
  const auto& run = sm.run[lo];
  const uint32_t chunk = run.bc & 0x1FFFFFFFu;
  const uint32_t page = (run.bc >> 29) * 64u + nth_set_bit(run.bits, r - run.r0);
  return chunk_addr(m, chunk) + kMeta + (uint64_t)page * (I::kSmallest << qi);
}

// Chunk-queue frees of one instance: bit, then the count per chunk (one add per chunk group of the
// wave), re-enqueue on the 0 -> >0 crossing; the class signal follows at block level.
template <class I>
__device__ inline void block_release_chunks(Memory& m, I& inst, ReuseRing& base, bool here, uint32_t qi,
                                            uint32_t chunk, uint32_t page) {
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[here ? qi : 0u];
  ChunkHdr* h = hdr<ChunkHdr>(m, here ? chunk : 0u);
  if (here) or_(&h->mask[page / 64u], 1ull << (page % 64u));
  order();
  const uint64_t grp = match<64>(here ? (uint64_t)chunk : ~0ull);
  const int L = first(grp);
  const uint32_t n = popc(grp);
  bool crossing = false;
  if (here && (int)lane() == L) {
    const int32_t old = add(&h->count, (int32_t)n);
    crossing = old <= 0 && old + (int32_t)n > 0;
  }
  ensure_in_queue<64>(x, cq, crossing, chunk);
}

}  // namespace detail
}  // namespace ouro
