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

// LDS scratch of one workgroup serving kTile requests per thread. Chunk queues write one page index
// (MemoryIndex) per served rank.
template <bool kChunks, uint32_t kTile = 1>
struct BlockScratch {
  static constexpr uint32_t kTileSize = kTile;
  uint32_t off[kTile * kBlockMaxWaves][kBlockMaxClasses];  // requests of (sub-tile, wave, class), then their rank offset
  uint32_t total[kBlockMaxClasses];
  uint32_t base[kBlockMaxClasses];    // page queues: first queue position; chunk queues: first pidx entry
  uint32_t served[kBlockMaxClasses];  // chunk queues: ranks served (the rest get nullptr)
  uint32_t pidx[kChunks ? 1024 * kTile : 1];  // chunk queues: page of each rank, class stretches in class order
};

namespace detail {

// Ranks of this thread's T requests among the block's requests of their classes gc[t] (kMarker: no
// request), after the class totals and per-(sub-tile, wave) offsets are in sm. Request t of thread i
// is request t * blockDim.x + i of the tile, and ranks follow that order: a block serving T requests
// per thread ranks exactly as a block of T * blockDim.x threads would. Ends with the block
// synchronised.
template <uint32_t T, class S>
__device__ inline void block_rank_tile(S& sm, const uint32_t (&gc)[T], uint32_t (&r)[T], uint32_t nclasses) {
  static_assert(T <= S::kTileSize, "scratch too small for the tile");
  const uint32_t wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63u) >> 6;
  uint32_t in[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    uint32_t* row = sm.off[t * kBlockMaxWaves + wave];
    if (lane() < kBlockMaxClasses) row[lane()] = 0;
    const bool valid = gc[t] != kMarker;
    const uint64_t grp = match<64>(valid ? (uint64_t)gc[t] : ~0ull);
    in[t] = rank_in(grp);
    if (valid && (int)lane() == first(grp)) row[gc[t]] = popc(grp);
  }
  __syncthreads();
  // (measured: one exclusive wave scan per class over the rows was no faster at four requests per
  // thread and 10% slower at one than this per-class walk by the class's lane of wave 0)
  if (wave == 0 && lane() < nclasses) {
    uint32_t run = 0;
    for (uint32_t t = 0; t < T; ++t)
      for (uint32_t w = 0; w < nwaves; ++w) {
        const uint32_t c = sm.off[t * kBlockMaxWaves + w][lane()];
        sm.off[t * kBlockMaxWaves + w][lane()] = run;
        run += c;
      }
    sm.total[lane()] = run;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) r[t] = gc[t] != kMarker ? sm.off[t * kBlockMaxWaves + wave][gc[t]] + in[t] : 0u;
}

template <class S>
__device__ inline uint32_t block_rank(S& sm, uint32_t gc, uint32_t nclasses) {
  const uint32_t g[1] = {gc};
  uint32_t r[1];
  block_rank_tile<1>(sm, g, r, nclasses);
  return r[0];
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
  // front_ positions and the semaphore claim in flight together (see alloc_pages)
  uint32_t pos = 0;
  if (lead) pos = add(&cq->front, total);
  auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_pages(x, lq, k); };
  sem_claim_n(x, cq, lead, total, carve);
  if (lead) sm.base[l] = pos;
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

// ---- standard page rings, a whole tile at once
// The T requests of a thread take (or return) their ring slots with all T loads in flight together;
// only a slot found not yet written (take) or not yet emptied (put) is waited for. A standard ring
// changes nothing but the slots, so the order of the T hand-overs does not matter. Each slot has one
// writer and one reader per lap (ring_slot_put / ring_slot_take).
template <uint32_t T, class I, class S>
__device__ inline void block_take_ring_tile(Memory& m, I& inst, uint32_t off, const S& sm, const uint32_t (&gc)[T],
                                            const uint32_t (&r)[T], void* (&out)[T]) {
  static_assert(!I::kChunks && is_std(I::kKind), "standard page rings only");
  uint32_t* slot[T];
  uint32_t val[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    slot[t] = nullptr;
    val[t] = kMarker;
    if (gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ) {
      slot[t] = &inst.tab.ring[gc[t] - off][(sm.base[gc[t]] + r[t]) & (I::kQueueEntries - 1)];
      val[t] = ld(slot[t]);
    }
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (!slot[t]) continue;
    if (val[t] == kMarker) val[t] = ring_slot_take(m, slot[t]);  // not yet published: wait for it
    else st(slot[t], kMarker);
    void* p = nullptr;
    const uint32_t qi = gc[t] - off;
    if (val[t] != kMarker && val[t] != kNullPage) {
      if (!chunk_ok(m, index_chunk(val[t])) || index_page(val[t]) >= ((I::kChunkSize >> I::kSmallLog2) >> qi))
        set_error(m, ERR_CORRUPT);
      else
        p = chunk_addr(m, index_chunk(val[t])) + kMeta + (uint64_t)index_page(val[t]) * (I::kSmallest << qi);
    }
    out[t] = p;
  }
}

// ---- VA page queues, a whole tile at once (take_lanes' VA branch for T positions)
// Phase A: the tile's pointer-ring lookups, queue-chunk checks and slot takes each in flight together;
// c[t] keeps the queue chunk each take emptied a slot of. Phase B (per sub-tile, in the order a
// sub-tile-by-sub-tile take would run it, so emptied queue chunks reach the reuse ring in the same
// order): the queue-chunk counts, one subtraction per queue chunk and wave, and the recycling of a
// queue chunk the last take emptied.
template <uint32_t T, class I, class S>
__device__ inline void block_take_va_tile_slots(Memory& m, I& inst, ReuseRing& base, uint32_t off, const S& sm,
                                                const uint32_t (&gc)[T], const uint32_t (&r)[T], void* (&out)[T],
                                                uint32_t (&c)[T]) {
  static_assert(!I::kChunks && is_va(I::kKind), "VA page queues only");
  Ctx<I> x{m, inst, base};
  bool here[T];
  uint32_t pos[T];
  ClassQueue* cq[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    here[t] = gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ;
    cq[t] = &inst.q[here[t] ? gc[t] - off : 0u];
    pos[t] = here[t] ? sm.base[gc[t]] + r[t] : 0u;
    c[t] = here[t] ? ld(&ring_of(x, cq[t])[(pos[t] / kQCSpots) & (I::kQueueEntries - 1)]) : kMarker;
  }
  uint32_t vs[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (here[t] && (c[t] == kMarker || is_qc_tag(c[t]))) c[t] = va_wait_qc(x, cq[t], pos[t]);
    vs[t] = here[t] && chunk_ok(m, c[t]) ? ld(&hdr<QueueChunkHdr>(m, c[t])->virtual_start) : 0u;
  }
  uint32_t* slot[T];
  uint32_t val[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (here[t] && (!chunk_ok(m, c[t]) || vs[t] / kQCSpots != pos[t] / kQCSpots)) {
      if (c[t] != kMarker) set_error(m, ERR_VSTART);
      c[t] = kMarker;
    }
    slot[t] = c[t] != kMarker ? &qc_slots(m, c[t])[pos[t] % kQCSpots] : nullptr;
    val[t] = slot[t] ? ld(slot[t]) : kMarker;
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    void* p = nullptr;
    if (slot[t]) {
      if (val[t] == kMarker) val[t] = ring_slot_take(m, slot[t]);  // not yet published: wait for it
      else st(slot[t], kMarker);
      const uint32_t qi = gc[t] - off;
      if (val[t] != kMarker && val[t] != kNullPage) {
        if (!chunk_ok(m, index_chunk(val[t])) || index_page(val[t]) >= ((I::kChunkSize >> I::kSmallLog2) >> qi))
          set_error(m, ERR_CORRUPT);
        else
          p = chunk_addr(m, index_chunk(val[t])) + kMeta + (uint64_t)index_page(val[t]) * (I::kSmallest << qi);
      }
    }
    if (here[t]) out[t] = p;
  }
}

// Phase B for sub-tile t of one instance (after an order() that completes every slot take).
template <uint32_t T, class I>
__device__ inline void block_take_va_tile_count(Memory& m, I& inst, ReuseRing& base, uint32_t off,
                                                const uint32_t (&gc)[T], uint32_t t, const uint32_t (&c)[T],
                                                uint32_t pos_qc) {
  const uint64_t grp = match<64>(c[t] != kMarker ? (uint64_t)c[t] : ~0ull);
  bool reclaim = false;
  uint32_t victim = 0;
  if (c[t] != kMarker && (int)lane() == first(grp)) {
    const uint32_t taken = popc(grp) << 16;
    const uint32_t old = sub(&hdr<QueueChunkHdr>(m, c[t])->count, taken);
    if (old - taken == kQCSpots) {  // QueueChunk::checkChunkEmptyDequeue
      Ctx<I> x{m, inst, base};
      ClassQueue* cq = &inst.q[gc[t] - off];
      victim = xchg(&ring_of(x, cq)[pos_qc & (I::kQueueEntries - 1)], qc_tag(pos_qc));
      reclaim = chunk_ok(m, victim);
    }
  }
  ring_put(m, base, reclaim, victim);
}

template <uint32_t T, class I, class S>
__device__ inline void block_put_ring_tile(Memory& m, I& inst, uint32_t off, const S& sm, const uint32_t (&gc)[T],
                                           const uint32_t (&r)[T], const uint32_t (&chunk)[T],
                                           const uint32_t (&page)[T]) {
  static_assert(!I::kChunks && is_std(I::kKind), "standard page rings only");
  uint32_t* slot[T];
  uint32_t cur[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    slot[t] = nullptr;
    cur[t] = kMarker;
    if (gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ) {
      slot[t] = &inst.tab.ring[gc[t] - off][(sm.base[gc[t]] + r[t]) & (I::kQueueEntries - 1)];
      cur[t] = ld(slot[t]);
    }
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (!slot[t]) continue;
    // a slot that still holds the entry of position - ring_size waits for its take
    if (cur[t] == kMarker) st(slot[t], make_index(chunk[t], page[t]));
    else ring_slot_put(m, slot[t], make_index(chunk[t], page[t]));
  }
  order();  // the slots are written before the block signals the pages
}

// put_lanes_after's VA branch for T positions per lane at once (after put_prealloc of all of them): the
// pointer-ring lookups, queue-chunk checks and slot writes each in flight together, then the
// queue-chunk counts per position in order. Page queues hand a slot over with a load and a store (one
// writer and one reader per position), chunk queues with a CAS on MARKER.
template <uint32_t T, class I>
__device__ inline void va_put_tile(Memory& m, I& inst, ReuseRing& base, ClassQueue* const (&cq)[T], const bool (&here)[T],
                                   const uint32_t (&pos)[T], const uint32_t (&val)[T]) {
  constexpr Kind K = I::kKind;
  static_assert(is_va(K), "VA queues only");
  Ctx<I> x{m, inst, base};
  uint32_t c[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t)
    c[t] = here[t] ? ld(&ring_of(x, cq[t])[(pos[t] / kQCSpots) & (I::kQueueEntries - 1)]) : kMarker;
  uint32_t vs[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (here[t] && (c[t] == kMarker || is_qc_tag(c[t]))) c[t] = va_wait_qc(x, cq[t], pos[t]);  // pre-allocation in flight
    vs[t] = here[t] && chunk_ok(m, c[t]) ? ld(&hdr<QueueChunkHdr>(m, c[t])->virtual_start) : 0u;
  }
  uint32_t* slot[T];
  uint32_t cur[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (here[t] && (!chunk_ok(m, c[t]) || vs[t] / kQCSpots != pos[t] / kQCSpots)) {
      if (c[t] != kMarker) set_error(m, ERR_VSTART);
      c[t] = kMarker;
    }
    slot[t] = c[t] != kMarker ? &qc_slots(m, c[t])[pos[t] % kQCSpots] : nullptr;
    if constexpr (is_chunk_kind(K)) cur[t] = slot[t] ? cas(slot[t], kMarker, val[t]) : kMarker;
    else cur[t] = slot[t] ? ld(slot[t]) : kMarker;
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (!slot[t]) continue;
    if constexpr (is_chunk_kind(K)) {
      if (cur[t] != kMarker) {
        Spin it(&m.error);
        while (cas(slot[t], kMarker, val[t]) != kMarker) {
          if (it.expired()) { OURO_TIMEOUT(m); break; }
          backoff(it.it);
        }
      }
    } else {
      // a slot still holding the entry of the previous lap waits for its take
      if (cur[t] == kMarker) st(slot[t], val[t]);
      else ring_slot_put(m, slot[t], val[t]);
    }
  }
  if constexpr (!is_chunk_kind(K)) order();  // plain slot stores complete before the counts
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint64_t grp = match<64>(c[t] != kMarker ? (uint64_t)c[t] : ~0ull);
    bool reclaim = false;
    uint32_t victim = 0;
    if (c[t] != kMarker && (int)lane() == first(grp)) {
      const uint32_t added = popc(grp) * 0x10001u;
      const uint32_t old = add(&hdr<QueueChunkHdr>(m, c[t])->count, added);
      if (qc_empty_after_enqueue(old, added)) {
        victim = xchg(&ring_of(x, cq[t])[(pos[t] / kQCSpots) & (I::kQueueEntries - 1)], qc_tag(pos[t] / kQCSpots));
        reclaim = chunk_ok(m, victim);
      }
    }
    ring_put(m, base, reclaim, victim);
  }
}

// Virtualized page queues (VA / VL): the tile's frees of one instance at the positions the block's
// back_ fetch-adds gave them. Every sub-tile's queue-chunk pre-allocation runs before any put waits: a
// put waits for the queue chunk (VL: the link to it) that the put at an earlier position allocates,
// and ranks put sub-tile t's positions below sub-tile t + 1's, so no put waits on a pre-allocation
// its own wave has still to make.
template <uint32_t T, class I, class S>
__device__ inline void block_put_pages_tile(Memory& m, I& inst, ReuseRing& base, uint32_t off, const S& sm,
                                            const uint32_t (&gc)[T], const uint32_t (&r)[T],
                                            const uint32_t (&chunk)[T], const uint32_t (&page)[T]) {
  static_assert(!I::kChunks && !is_std(I::kKind), "virtualized page queues only");
  Ctx<I> x{m, inst, base};
  bool here[T];
  uint32_t pos[T], qc[T];
  ClassQueue* cq[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    here[t] = gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ;
    cq[t] = &inst.q[here[t] ? gc[t] - off : 0u];
    pos[t] = here[t] ? sm.base[gc[t]] + r[t] : 0u;
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) qc[t] = put_prealloc(x, cq[t], here[t], pos[t]);
  if constexpr (is_va(I::kKind)) {
    uint32_t val[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) val[t] = make_index(chunk[t], page[t]);
    va_put_tile<T>(m, inst, base, cq, here, pos, val);
  } else {
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) put_lanes_after<64>(x, cq[t], here[t], pos[t], make_index(chunk[t], page[t]), qc[t]);
  }
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
// A class's block total is claimed by one whole wave (class gc goes to wave gc mod waves, so the
// classes of a block proceed in parallel). Each step looks at up to 64 * kClaimPer queue positions,
// kClaimPer consecutive ones per lane, all slot and header loads in flight together (measured: with 2
// or 4 per lane the persistent kernels spill registers and the claims got slower, so 1): a prefix sum over
// the chunks' free counts (in position order) divides the pages, each lane takes its share of each of
// its chunks with one count CAS, removes a chunk it emptied at the position it found it at, and claims
// the bits from the chunks' 64-bit mask words (the first word of every chunk in flight together),
// writing the page of each rank it serves into sm.pidx. front_ is compacted once, after the claim, if
// it removed anything. The reference walks the queue one chunk at a time per thread
// (ChunkQueue_impl.cuh:88-151); a serial walk per class leader leaves a block waiting on its largest
// class, where every chunk holds a single page.
constexpr uint32_t kClaimPer = 1;

template <class I, class S>
__device__ inline void block_claim_chunks(Memory& m, I& inst, ReuseRing& base, uint32_t off, S& sm,
                                          uint32_t nclasses) {
  constexpr Kind K = I::kKind;
  constexpr uint32_t P = kClaimPer;
  const uint32_t wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63u) >> 6;
  const uint32_t l = lane();
  Ctx<I> x{m, inst, base};
  for (uint32_t gc = off; gc < off + I::kNQ; ++gc) {
    if (gc % nwaves != wave) continue;
    const uint32_t total = sm.total[gc];
    if (total == 0) continue;
    ClassQueue* cq = &inst.q[gc - off];
    uint32_t start = 0;
    for (uint32_t c = 0; c < gc; ++c) start += sm.total[c];
    for (uint32_t i = l; i < total; i += kWave) sm.pidx[start + i] = kMarker;
    auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_chunks(x, lq, k); };
    sem_claim_n(x, cq, l == 0, total, carve);
    const uint32_t psize = page_size_of(x, cq);
    uint32_t v = 0;
    if (l == 0) v = draw_start(x, cq, total);
    v = uniform(v);
    // (no redraws here: a whole wave passes 64 * kClaimPer positions a step, and measured, drawing
    // new starts scattered the block claims' consumption order and doubled their kernel time)
    uint32_t assigned = 0, wraps = 0, cursor = kMarker, reclaim = kMarker;
    bool any_removed = false;
    Spin it(&m.error);
    while (assigned < total) {
      uint32_t b = 0, nb = 0;
      bool give_up = false;
      if (l == 0) {
        b = ld(&cq->back);
        wraps += scan_wrap<I>(cq, v, cursor, b);
        nb = b - v;  // 64 * kClaimPer positions at a time, not past back_, inside one queue chunk for VA / VL
        if (nb >= 0x80000000u) nb = 0;
        if (nb > kWave * P) nb = kWave * P;
        if constexpr (!is_std(K)) nb = nb < kQCSpots - v % kQCSpots ? nb : kQCSpots - v % kQCSpots;
        give_up = (wraps >= 2 || nb == 0) && pool_exhausted(m);
      }
      v = uniform(v);
      cursor = uniform(cursor);
      wraps = uniform(wraps);
      nb = uniform(nb);
      if (ballot(give_up)) {
        // the pool ran out and a whole window pass found nothing: unserved ranks get nullptr
        if (l == 0) sem_signal(cq, total - assigned);
        break;
      }
      if (nb == 0) {
        if (ballot(it.expired())) {
          OURO_TIMEOUT(m);
          break;
        }
        backoff(it.it);
        continue;
      }
      if (l == 0) {
        OURO_DIAG_ADD(0, 1);
        OURO_DIAG_ADD(2, nb);
      }
      // the slots of positions v .. v + nb (one queue chunk for VA / VL): lane l holds v + l * P + j
      uint32_t s[P], qc = kMarker;
      uint32_t* slot[P];
      const uint32_t* sl = nullptr;
      if constexpr (!is_std(K)) {
        uint32_t fill = kMarker, dead = 0;
        if (l == 0) sl = slots_at(x, cq, v, cursor, fill, &qc, &dead);
        sl = reinterpret_cast<const uint32_t*>(readlane64(reinterpret_cast<uint64_t>(sl), 0));
        cursor = uniform(cursor);
        fill = uniform(fill);
        qc = uniform(qc);
        dead = uniform(dead);
        if (dead) {
          // the queue chunk holds no live entry: pass the rest of it in one step
          v += dead;
          if (ballot(it.expired())) {
            OURO_TIMEOUT(m);
            break;
          }
          continue;
        }
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
          const uint32_t i = l * P + j;
          slot[j] = const_cast<uint32_t*>(sl) + i;
          s[j] = i < nb ? (sl ? ld(&sl[i]) : fill) : kMarker;
        }
      } else {
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
          const uint32_t i = l * P + j;
          slot[j] = &ring_of(x, cq)[(v + i) & (I::kQueueEntries - 1)];
          s[j] = i < nb ? ld(slot[j]) : kMarker;
        }
      }
      // the chunks' headers, all in flight
      uint64_t cg[P], qp[P];
      uint32_t ps[P], sz[P];
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        ChunkHdr* h = hdr<ChunkHdr>(m, chunk_ok(m, s[j]) ? s[j] : 0u);
        cg[j] = ld(count_gen(h));
        qp[j] = ld(&h->queue_pos);
        ps[j] = ld(&h->page_size);
        sz[j] = ld(&h->size);
      }
      int32_t cnt[P];
      uint32_t g[P], c[P], lane_sum = 0;
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        cnt[j] = cg_count(cg[j]);
        g[j] = cg_gen(cg[j]);
        const bool ok = chunk_ok(m, s[j]) && cnt[j] > 0 && qp[j] == qp_at(g[j], v + l * P + j) && ps[j] == psize &&
                        gen_class(g[j]) == gc - off;
        c[j] = ok ? (uint32_t)cnt[j] : 0u;
        lane_sum += c[j];
      }
      uint32_t offered = 0;
      uint32_t before = wave_exclusive_sum(lane_sum, &offered);
      const uint32_t rem = total - assigned;
      uint32_t t[P], lane_took = 0;
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        // share = clamp(rem - before, 0, c), written branch-free: with the bound inside an `if` the
        // ROCm 7.2 compiler hoisted the min out of the branch and dropped the before < rem guard
        const uint32_t owed = before < rem ? rem - before : 0u;
        t[j] = owed < c[j] ? owed : c[j];
        before += c[j];
      }
      // changed (or recycled) meanwhile: skip it (the CASes in flight together)
      uint64_t seen[P];
#pragma unroll
      for (uint32_t j = 0; j < P; ++j)
        seen[j] = t[j] ? cas(count_gen(hdr<ChunkHdr>(m, s[j])), make_cg(cnt[j], g[j]), make_cg(cnt[j] - (int32_t)t[j], g[j]))
                       : 0ull;
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        if (t[j] && seen[j] != make_cg(cnt[j], g[j])) t[j] = 0;
        lane_took += t[j];
      }
      uint32_t took = 0;
      const uint32_t r0 = assigned + wave_exclusive_sum(lane_took, &took);
      // chunks this claim emptied leave the queue at the position they were found at (cq_remove for
      // the whole batch: slots tombed in parallel, one count update per queue and queue chunk)
      bool removed[P];
      uint32_t lane_rm = 0;
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        const uint64_t want_qp = qp_at(g[j], v + l * P + j);
        removed[j] = t[j] && t[j] == c[j] && cas(&hdr<ChunkHdr>(m, s[j])->queue_pos, want_qp, qp_not(g[j])) == want_qp;
        lane_rm += removed[j] ? 1u : 0u;
      }
      uint32_t nrm = 0;
      (void)wave_exclusive_sum(lane_rm, &nrm);
      if (l == 0) OURO_DIAG_ADD(8, nrm);
      if (nrm) {
        any_removed = true;
#pragma unroll
        for (uint32_t j = 0; j < P; ++j)
          if (removed[j]) xchg(slot[j], kTomb);
        uint32_t rc = kMarker;
        if (l == 0) {
          sub(&cq->count, (int32_t)nrm);
          if constexpr (!is_std(K)) {
            const uint32_t old = sub(&hdr<QueueChunkHdr>(m, qc)->count, nrm << 16);
            if (old - (nrm << 16) == kQCSpots) {  // every position of the queue chunk is dequeued
              if constexpr (is_va(K)) {
                const uint32_t victim = xchg(&ring_of(x, cq)[(v / kQCSpots) & (I::kQueueEntries - 1)], qc_tag(v / kQCSpots));
                if (chunk_ok(m, victim)) rc = victim;
              } else {
                rc = vl_set_front(x, cq, qc) | 0x80000000u;
              }
            }
          }
        }
        rc = uniform(rc);
        if (rc != kMarker) {
          if (reclaim != kMarker) finish_reclaim(x, cq, l == 0, reclaim);
          reclaim = rc;
        }
        order();
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
          // a free raced in
          const bool again = removed[j] && ld(&hdr<ChunkHdr>(m, s[j])->count) > 0;
          if (ballot(again)) ensure_in_queue<1>(x, cq, again, s[j], g[j]);
        }
      }
      // mask claims: the first word of every chunk with pages to take, all in flight; a chunk whose
      // pages run over into further words continues word by word
      uint32_t w[P], left[P], r = start + r0;
      uint64_t mw[P];
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        const uint32_t nw = mask_words(sz[j]);
        w[j] = t[j] ? ((sz[j] - (uint32_t)cnt[j]) / 64u) % nw : 0u;
        left[j] = t[j];
        mw[j] = t[j] ? ld(&hdr<ChunkHdr>(m, s[j])->mask[w[j]]) : 0ull;
      }
      uint64_t take[P], got[P];
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        take[j] = left[j] ? lowest_bits(mw[j], left[j]) : 0ull;
        got[j] = take[j] ? and_(&hdr<ChunkHdr>(m, s[j])->mask[w[j]], ~take[j]) & take[j] : 0ull;
      }
#pragma unroll
      for (uint32_t j = 0; j < P; ++j) {
        if (!t[j]) continue;
        const uint32_t nw = mask_words(sz[j]);
        ChunkHdr* h = hdr<ChunkHdr>(m, s[j]);
        uint64_t gb = got[j];
        bool whole = take[j] != 0 && gb == take[j];
        if (take[j] == 0) w[j] = (w[j] + 1) % nw;
        Spin bits_wait(&m.error);
        while (true) {
          left[j] -= popc(gb);
          while (gb) {
            const uint32_t bit = (uint32_t)__builtin_ctzll(gb);
            gb &= gb - 1;
            sm.pidx[r++] = make_index(s[j], w[j] * 64u + bit);
          }
          if (left[j] == 0) break;
          if (whole) w[j] = (w[j] + 1) % nw;
          if (bits_wait.expired()) {
            OURO_TIMEOUT(m);
            break;
          }
          const uint64_t tk = lowest_bits(ld(&h->mask[w[j]]), left[j]);
          if (tk == 0) {
            w[j] = (w[j] + 1) % nw;
            gb = 0;
            whole = false;
            continue;
          }
          gb = and_(&h->mask[w[j]], ~tk) & tk;
          whole = gb == tk;
        }
      }
      assigned += took;
      v += nb;
      if (assigned < total && ballot(it.expired())) {
        OURO_TIMEOUT(m);
        break;
      }
    }
    // front_ over the leading positions this claim emptied (one wave-wide compaction per claim). A
    // standard ring's compaction loads, passes and resets every dead position (a VA / VL one passes an
    // emptied queue chunk in one move), and after the block claims of a round it is what the slowest
    // waves wait on: 64 positions per lane and round there (CQ malloc, 1M mixed requests, tiles 1 / 2 /
    // 4: 0.84 / 0.76 / 1.12 ms with 16, 0.73 / 0.75 / 0.94 with 32, 0.82 / 0.62 / 0.71 with 64; the
    // per-thread paths and the VA / VL kinds measured slower with wider rounds).
    if (ballot(any_removed)) compact_wave<is_std(K) ? 64u : 16u>(x, cq);
    if (l == 0) {
      sm.base[gc] = start;
      sm.served[gc] = assigned;
    }
    finish_reclaim(x, cq, l == 0 && reclaim != kMarker, reclaim);
  }
  (void)nclasses;
}

// Page of rank r of class gc, from the claim's page list.
template <class I, class S>
__device__ inline void* block_page_of(Memory& m, I& inst, const S& sm, bool here, uint32_t gc, uint32_t qi, uint32_t r) {
  if (!here || r >= sm.served[gc]) return nullptr;
  const uint32_t idx = sm.pidx[sm.base[gc] + r];
  if (idx == kMarker) return nullptr;
  return chunk_addr(m, index_chunk(idx)) + kMeta + (uint64_t)index_page(idx) * (I::kSmallest << qi);
}

// Chunk-queue frees of one instance: bit, then the count per chunk (one add per chunk group of the
// wave), recycling of chunks made full, re-enqueue on the 0 -> >0 crossing; the class signal follows
// at block level, less the pages that left with recycled chunks (taken off the block's class total
// in LDS, `total`).
//
// With `defer` (recycling off) a chunk whose count this release moved from <= 0 to > 0 is not enqueued
// here: the lane claims the enqueue (queue_pos not queued -> pending) and returns true with the chunk's
// generation in gen_out, and the workgroup enqueues every such chunk of the tile together afterwards
// (block_enqueue_back, block_enqueue_chunks).
template <class I>
__device__ inline bool block_release_chunks(Memory& m, I& inst, ReuseRing& base, bool here, uint32_t qi,
                                            uint32_t chunk, uint32_t page, uint32_t* total, bool defer,
                                            uint32_t& gen_out) {
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[here ? qi : 0u];
  ChunkHdr* h = hdr<ChunkHdr>(m, here ? chunk : 0u);
  // the bits first (ChunkAccess::freePage, ChunkAccess_impl.cuh:7-25): one OR per mask word the wave
  // frees pages of, carrying all of them (a tile frees the pages its malloc took together, so its
  // lanes share words), instead of one per page
  {
    const uint64_t key = here ? ((uint64_t)chunk << 3) | (page / 64u) : ~0ull;
    for (uint64_t todo = ballot(here); todo;) {
      const int l = first(todo);
      const uint64_t same = ballot(key == readlane64(key, l)) & todo;
      todo &= ~same;
      uint64_t bits = 0;
      for (uint64_t r = same; r; r &= r - 1) bits |= 1ull << (readlane(page, first(r)) % 64u);
      if ((int)lane() == l) or_(&h->mask[page / 64u], bits);
    }
  }
  order();
  const uint64_t grp = match<64>(here ? (uint64_t)chunk : ~0ull);
  const int L = first(grp);
  const uint32_t n = popc(grp);
  const bool recycle = m.recycle_min != kRecycleOff;
  bool crossing = false, full = false;
  uint32_t gen = 0, size = 0;
  if (here && (int)lane() == L) {
    if (recycle) {  // read while the group's pages still pin the chunk's life
      gen = ld(&h->gen);
      size = ld(&h->size);
    }
    const int32_t old = add(&h->count, (int32_t)n);
    crossing = old <= 0 && old + (int32_t)n > 0;
    if (!recycle && crossing) gen = ld(&h->gen);
    full = recycle && old + (int32_t)n == (int32_t)size;
  }
  const bool gone = flash_chunks(x, cq, full, chunk, gen, size, size - n, crossing ? 1u : 0u);
  if (gone) __hip_atomic_fetch_sub(total, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (defer) {
    const bool won = crossing && !gone && cas(&h->queue_pos, qp_not(gen), qp_pending(gen)) == qp_not(gen);
    if (won) gen_out = gen;  // lanes of other instances keep theirs
    return won;
  }
  ensure_in_queue<64>(x, cq, crossing && !gone, chunk, gen);
  return false;
}

// block_release_chunks for a whole tile with the enqueues deferred (recycling off): the tile's bit
// ORs in flight together, one order(), then the count adds of every sub-tile in flight, then the
// enqueue claims. won[t] (gen_out[t]) as block_release_chunks returns them, for the lanes whose request
// t belongs to this instance.
template <uint32_t T, class I>
__device__ inline void block_release_chunks_tile(Memory& m, I& inst, uint32_t off, const uint32_t (&gc)[T],
                                                 const uint32_t (&chunk)[T], const uint32_t (&page)[T],
                                                 bool (&won)[T], uint32_t (&gen_out)[T]) {
  static_assert(I::kChunks, "chunk queues only");
  bool here[T];
  ChunkHdr* h[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    here[t] = gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ;
    h[t] = hdr<ChunkHdr>(m, here[t] ? chunk[t] : 0u);
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint64_t key = here[t] ? ((uint64_t)chunk[t] << 3) | (page[t] / 64u) : ~0ull;
    for (uint64_t todo = ballot(here[t]); todo;) {
      const int l = first(todo);
      const uint64_t same = ballot(key == readlane64(key, l)) & todo;
      todo &= ~same;
      uint64_t bits = 0;
      for (uint64_t r = same; r; r &= r - 1) bits |= 1ull << (readlane(page[t], first(r)) % 64u);
      if ((int)lane() == l) or_(&h[t]->mask[page[t] / 64u], bits);
    }
  }
  order();
  bool lead[T];
  int32_t old[T], n[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint64_t grp = match<64>(here[t] ? (uint64_t)chunk[t] : ~0ull);
    lead[t] = here[t] && (int)lane() == first(grp);
    n[t] = (int32_t)popc(grp);
    old[t] = lead[t] ? add(&h[t]->count, n[t]) : 1;
  }
  uint32_t gen[T];
  bool crossing[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    crossing[t] = lead[t] && old[t] <= 0 && old[t] + n[t] > 0;
    gen[t] = crossing[t] ? ld(&h[t]->gen) : 0u;
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const bool w = crossing[t] && cas(&h[t]->queue_pos, qp_not(gen[t]), qp_pending(gen[t])) == qp_not(gen[t]);
    if (here[t]) {
      won[t] = w;
      if (w) gen_out[t] = gen[t];
    }
  }
}

// Wave 0, once the tile's deferred enqueues are ranked per class (sm.total, sm.off): one count_ and
// one back_ fetch-add per class (ChunkQueue::enqueue, ChunkQueue_impl.cuh:20-38, for the whole tile).
template <class I, class S>
__device__ inline void block_enqueue_back(I& inst, uint32_t off, S& sm) {
  static_assert(I::kChunks, "chunk queues only");
  const uint32_t l = lane();
  const bool mine = l >= off && l < off + I::kNQ;
  const uint32_t total = mine ? sm.total[l] : 0u;
  if (total) {
    add(&inst.q[l - off].count, (int32_t)total);
    sm.base[l] = add(&inst.q[l - off].back, total);
  }
}

// The tile's deferred enqueues of one instance: chunk[t] goes to position sm.base[ec[t]] + er[t]. Every
// queue-chunk pre-allocation of the T positions runs before any put waits (a put may wait for a queue
// chunk that a later position of another workgroup's tile pre-allocates); queue_pos names the position
// once the put is complete.
template <uint32_t T, class I, class S>
__device__ inline void block_enqueue_chunks(Memory& m, I& inst, ReuseRing& base, uint32_t off, const S& sm,
                                            const uint32_t (&ec)[T], const uint32_t (&er)[T],
                                            const uint32_t (&chunk)[T], const uint32_t (&gen)[T]) {
  static_assert(I::kChunks, "chunk queues only");
  Ctx<I> x{m, inst, base};
  bool here[T];
  uint32_t pos[T], qc[T];
  ClassQueue* cq[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    here[t] = ec[t] != kMarker && ec[t] >= off && ec[t] < off + I::kNQ;
    cq[t] = &inst.q[here[t] ? ec[t] - off : 0u];
    pos[t] = here[t] ? sm.base[ec[t]] + er[t] : 0u;
  }
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) qc[t] = put_prealloc(x, cq[t], here[t], pos[t]);
  if constexpr (is_va(I::kKind)) {
    va_put_tile<T>(m, inst, base, cq, here, pos, chunk);
  } else {
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) put_lanes_after<64>(x, cq[t], here[t], pos[t], chunk[t], qc[t]);
  }
  order();
#pragma unroll
  for (uint32_t t = 0; t < T; ++t)
    if (here[t]) st(&hdr<ChunkHdr>(m, chunk[t])->queue_pos, qp_at(gen[t], pos[t]));
}

}  // namespace detail
}  // namespace ouro
