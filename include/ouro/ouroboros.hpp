// Ouroboros for MI355X: the public device API (malloc / free from any kernel thread) and its host
// half (initialize / getDeviceMemoryManager), header-only like the reference.
//
// Reference surface kept: Ouroboros<Instances...> with
//   void  initialize(size_t instantiation_size, size_t additionalSizeBeginning = 0,
//                    size_t additionalSizeEnd = 0)                   (device/Ouroboros.cuh:281)
//   MyType* getDeviceMemoryManager()                                  (device/Ouroboros.cuh:297)
//   __device__ void* malloc(size_t size)                              (device/Ouroboros_impl.cuh:83-91)
//   __device__ void  free(void* ptr)                                  (device/Ouroboros_impl.cuh:95-113)
//   __device__ void  enqueueInitialChunk(queue, chunk, avail, ppc)    (device/Ouroboros.cuh:291-295)
// malloc/free take an optional aggregation width template argument (1, 16 or 64); the default
// OURO_DEFAULT_AGG is 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <tuple>
#include <utility>
#include <vector>

#include "detail/device.hpp"
#include "detail/layout.hpp"
#include "detail/instance.hpp"
#include "detail/block.hpp"

#ifndef OURO_DEFAULT_AGG
#define OURO_DEFAULT_AGG 64
#endif

namespace ouro {

// ------------------------------------------------------------------ instance chain storage
template <class... Is>
struct Chain;
template <>
struct Chain<> {};
template <class I, class... Rest>
struct Chain<I, Rest...> {
  I inst;
  Chain<Rest...> next;
};

namespace detail {

template <int W, class I>
__device__ inline void* inst_malloc(Memory& m, I& inst, ReuseRing& base, uint64_t size, bool want) {
  Ctx<I> x{m, inst, base};
  const uint32_t qi = want ? class_of(size, I::kSmallLog2) : 0u;
  if constexpr (I::kChunks) return alloc_chunk_pages<W>(x, qi, want);
  else return alloc_pages<W>(x, qi, want);
}

template <int W, class I>
__device__ inline void inst_free(Memory& m, I& inst, ReuseRing& base, uint32_t qi, bool want, uint32_t chunk,
                                 uint32_t page) {
  Ctx<I> x{m, inst, base};
  if constexpr (I::kChunks) free_chunk_pages<W>(x, qi, want, chunk, page);
  else free_pages<W>(x, qi, want, chunk, page);
}

// AllocationHelper::getNextPow2 (Helper.cuh:12-18).
__device__ __forceinline__ uint64_t next_pow2(uint64_t v) { return v <= 1 ? 1 : 1ull << (64 - __builtin_clzll(v - 1)); }

// Above the largest page of the chain: the device heap at the next power of two (the chain's end,
// Ouroboros<>::malloc, Ouroboros.cuh:398-408). nullptr and ERR_TOO_LARGE when the heap has none.
__device__ inline void* heap_malloc(Memory& m, uint64_t size) {
  void* p = ::malloc(next_pow2(size));
  if (p == nullptr) set_error(m, ERR_TOO_LARGE);
  return p;
}

// Pointers outside the whole pool allocation came from the device heap (Ouroboros_impl.cuh:98-104).
__device__ __forceinline__ bool in_pool(const Memory& m, const void* p) {
  const uint8_t* q = reinterpret_cast<const uint8_t*>(p);
  return q >= m.d_memory && q < m.d_memory + m.pool_bytes;
}

// Walk the chain: lanes whose size fits instance I allocate there, the rest move on.
template <int W, class I, class... Rest>
__device__ inline void* chain_malloc(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, uint64_t size, bool want) {
  const bool here = want && size <= I::kLargest;
  void* r = nullptr;
  if (dev::ballot(here)) {
    stat_wave(m, STAT_PAGE_ALLOC, here);
    void* p = inst_malloc<W>(m, ch.inst, base, size, here);
    if (here) r = p;
  }
  if constexpr (sizeof...(Rest) > 0) {
    const bool later = want && !here;
    if (dev::ballot(later)) {
      void* p = chain_malloc<W>(m, ch.next, base, size, later);
      if (later) r = p;
    }
  } else {
    if (want && !here) r = heap_malloc(m, size);
  }
  return r;
}

// freePageRecursive (Ouroboros_impl.cuh:117-125): route by the chunk's page size.
template <int W, class I, class... Rest>
__device__ inline void chain_free(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, uint32_t page_size,
                                  bool want, uint32_t chunk, uint64_t offset) {
  const bool here = want && page_size <= I::kLargest;
  if (dev::ballot(here)) {
    uint32_t qi = 0, page = 0;
    bool ok = here;
    if (here) {
      // page sizes are powers of two between kSmallest and kLargest
      ok = page_size >= I::kSmallest && (page_size & (page_size - 1)) == 0;
      if (ok) {
        qi = log2_floor(page_size) - I::kSmallLog2;
        page = (uint32_t)(offset >> log2_floor(page_size));
        ok = (offset & (page_size - 1)) == 0 && page < I::kChunkSize / page_size;
      }
      if (!ok) set_error(m, ERR_BAD_FREE);
    }
    stat_wave(m, STAT_PAGE_FREE, ok);
    inst_free<W>(m, ch.inst, base, qi, ok, chunk, page);
  }
  if constexpr (sizeof...(Rest) > 0) {
    const bool later = want && !here;
    if (dev::ballot(later)) chain_free<W>(m, ch.next, base, page_size, later, chunk, offset);
  } else {
    if (want && !here) set_error(m, ERR_BAD_FREE);
  }
}

// validOuroborosPointer (Ouroboros.cuh:350-356), tightened to the chunk region, and the chunk /
// page-size lookup of Ouroboros::free (Ouroboros_impl.cuh:95-113). A pointer outside the pool goes
// back to the device heap (Ouroboros_impl.cuh:98-104); one inside the pool that is not a block of it
// sets ERR_BAD_FREE.
__device__ inline bool route_free(Memory& m, const void* ptr, uint32_t& chunk, uint32_t& page_size, uint64_t& offset) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(ptr);
  if (p == nullptr) return false;
  if (!in_pool(m, p)) {
    ::free(const_cast<uint8_t*>(p));
    return false;
  }
  bool want = p >= m.d_data && p < m.d_data + (uint64_t)m.max_chunks * kStride;
  if (want) {
    const uint64_t diff = (uint64_t)(p - m.d_data);
    const uint32_t idx = (uint32_t)(diff / kStride);
    // the chunk's class byte, in place of ChunkLocator::getChunkIndex and CommonChunk::page_size:
    // without recycling a byte only ever changes from 0 to its final value, so a plain (L2-cached)
    // load is exact unless it reads 0, which is confirmed with a device-coherent load; a recycled
    // chunk changes class, so with recycling on the byte is always read coherently
    uint32_t b = m.recycle_min == kRecycleOff ? m.chunk_log2[idx] : 0u;
    if (b == 0) b = dev::ld(&m.chunk_log2[idx]);
    want = b != 0 && (b >> 5) <= idx;
    if (want) {
      const uint32_t c = idx - (b >> 5);
      chunk = c;
      page_size = 1u << (b & 31u);
      const uint64_t base = (uint64_t)c * kStride + kMeta;
      want = diff >= base;
      offset = diff - base;
    }
  }
  if (!want) set_error(m, ERR_BAD_FREE);
  return want;
}

// ---- block-wide calls over the chain: classes are numbered across instances (offset `off`)
template <class I, class... Rest>
__device__ inline uint32_t chain_alloc_class(uint64_t size, uint32_t off = 0) {
  if (size <= I::kLargest) return off + class_of(size, I::kSmallLog2);
  if constexpr (sizeof...(Rest) > 0) return chain_alloc_class<Rest...>(size, off + I::kNQ);
  else return kMarker;
}

// freePageRecursive (Ouroboros_impl.cuh:117-125): the instance whose page sizes include page_size.
template <class I, class... Rest>
__device__ inline uint32_t chain_free_class(uint32_t page_size, uint64_t offset, uint32_t& page, uint32_t off = 0) {
  if (page_size <= I::kLargest) {
    const bool ok = page_size >= I::kSmallest && (page_size & (page_size - 1)) == 0 &&
                    (offset & (page_size - 1)) == 0 && (offset >> log2_floor(page_size)) < I::kChunkSize / page_size;
    if (!ok) return kMarker;
    page = (uint32_t)(offset >> log2_floor(page_size));
    return off + log2_floor(page_size) - I::kSmallLog2;
  }
  if constexpr (sizeof...(Rest) > 0) return chain_free_class<Rest...>(page_size, offset, page, off + I::kNQ);
  else return kMarker;
}

template <class S, class I, class... Rest>
__device__ inline void chain_block_claim(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm, uint32_t nclasses,
                                         uint32_t off = 0) {
  // chunk queues: every wave claims its share of the classes; page queues: wave 0's class lanes
  if constexpr (I::kChunks) block_claim_chunks(m, ch.inst, base, off, sm, nclasses);
  else if (threadIdx.x < kWave) block_claim_pages(m, ch.inst, base, off, sm);
  if constexpr (sizeof...(Rest) > 0) chain_block_claim(m, ch.next, base, sm, nclasses, off + I::kNQ);
}

template <class S, class I, class... Rest>
__device__ inline void* chain_block_take(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm, uint32_t gc,
                                         uint32_t r, uint32_t off = 0) {
  const bool here = gc != kMarker && gc >= off && gc < off + I::kNQ;
  void* p = nullptr;
  if (dev::ballot(here)) {
    const uint32_t qi = here ? gc - off : 0u;
    void* q;
    if constexpr (I::kChunks) q = block_page_of(m, ch.inst, sm, here, here ? gc : 0u, qi, r);
    else q = block_take_page(m, ch.inst, base, here, qi, here ? sm.base[gc] + r : 0u);
    if (here) p = q;
  }
  if constexpr (sizeof...(Rest) > 0) {
    void* q = chain_block_take(m, ch.next, base, sm, gc, r, off + I::kNQ);
    if (!here && q) p = q;
  }
  return p;
}

// Whole-tile takes / puts of chains whose instances all use standard page rings.
template <uint32_t T, class S, class I, class... Rest>
__device__ inline void chain_take_ring_tile(Memory& m, Chain<I, Rest...>& ch, S& sm, const uint32_t (&gc)[T],
                                            const uint32_t (&r)[T], void* (&out)[T], uint32_t off = 0) {
  block_take_ring_tile<T>(m, ch.inst, off, sm, gc, r, out);
  if constexpr (sizeof...(Rest) > 0) chain_take_ring_tile<T>(m, ch.next, sm, gc, r, out, off + I::kNQ);
}

template <uint32_t T, class S, class I, class... Rest>
__device__ inline void chain_put_ring_tile(Memory& m, Chain<I, Rest...>& ch, S& sm, const uint32_t (&gc)[T],
                                           const uint32_t (&r)[T], const uint32_t (&chunk)[T],
                                           const uint32_t (&page)[T], uint32_t off = 0) {
  block_put_ring_tile<T>(m, ch.inst, off, sm, gc, r, chunk, page);
  if constexpr (sizeof...(Rest) > 0) chain_put_ring_tile<T>(m, ch.next, sm, gc, r, chunk, page, off + I::kNQ);
}

// Whole-tile takes of chains whose instances are all VA page queues: every instance's slot takes
// first (c[i] keeps instance i's queue chunks), then, after the caller's order(), the queue-chunk counts
// sub-tile by sub-tile and instance by instance, the order of a take per sub-tile.
template <uint32_t T, class S, class I, class... Rest>
__device__ __forceinline__ void chain_take_va_slots(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm,
                                                    const uint32_t (&gc)[T], const uint32_t (&r)[T], void* (&out)[T],
                                                    uint32_t (*c)[T], uint32_t off = 0) {
  block_take_va_tile_slots<T>(m, ch.inst, base, off, sm, gc, r, out, c[0]);
  if constexpr (sizeof...(Rest) > 0) chain_take_va_slots<T>(m, ch.next, base, sm, gc, r, out, c + 1, off + I::kNQ);
}
template <uint32_t T, class S, class I, class... Rest>
__device__ __forceinline__ void chain_take_va_count(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm,
                                                    const uint32_t (&gc)[T], const uint32_t (&r)[T], uint32_t t,
                                                    uint32_t (*c)[T], uint32_t off = 0) {
  const bool here = gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ;
  const uint32_t pq = here ? (sm.base[gc[t]] + r[t]) / kQCSpots : 0u;
  if (dev::ballot(c[0][t] != kMarker)) block_take_va_tile_count<T>(m, ch.inst, base, off, gc, t, c[0], pq);
  if constexpr (sizeof...(Rest) > 0) chain_take_va_count<T>(m, ch.next, base, sm, gc, r, t, c + 1, off + I::kNQ);
}

// Whole-tile chunk frees of chains of chunk queues with deferred enqueues (block_release_chunks_tile),
// instance by instance.
template <uint32_t T, class I, class... Rest>
__device__ __forceinline__ void chain_release_chunks_tile(Memory& m, Chain<I, Rest...>& ch, const uint32_t (&gc)[T],
                                                          const uint32_t (&chunk)[T], const uint32_t (&page)[T],
                                                          bool (&won)[T], uint32_t (&gen)[T], uint32_t off = 0) {
  bool any = false;
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) any |= gc[t] != kMarker && gc[t] >= off && gc[t] < off + I::kNQ;
  if (dev::ballot(any)) block_release_chunks_tile<T>(m, ch.inst, off, gc, chunk, page, won, gen);
  if constexpr (sizeof...(Rest) > 0) chain_release_chunks_tile<T>(m, ch.next, gc, chunk, page, won, gen, off + I::kNQ);
}

// Whole-tile puts of chains of page queues, standard rings or virtualized (instance by instance, in
// chain order on every workgroup).
template <uint32_t T, class S, class I, class... Rest>
__device__ inline void chain_put_pages_tile(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm,
                                            const uint32_t (&gc)[T], const uint32_t (&r)[T],
                                            const uint32_t (&chunk)[T], const uint32_t (&page)[T], uint32_t off = 0) {
  static_assert(!I::kChunks, "page queues only");
  if constexpr (is_std(I::kKind)) block_put_ring_tile<T>(m, ch.inst, off, sm, gc, r, chunk, page);
  else block_put_pages_tile<T>(m, ch.inst, base, off, sm, gc, r, chunk, page);
  if constexpr (sizeof...(Rest) > 0) chain_put_pages_tile<T>(m, ch.next, base, sm, gc, r, chunk, page, off + I::kNQ);
}

template <class S, class I, class... Rest>
__device__ inline void chain_block_back(Chain<I, Rest...>& ch, S& sm, uint32_t off = 0) {
  if constexpr (!I::kChunks) block_back(ch.inst, off, sm);
  if constexpr (sizeof...(Rest) > 0) chain_block_back(ch.next, sm, off + I::kNQ);
}

template <class S, class I, class... Rest>
__device__ inline void chain_block_signal(Chain<I, Rest...>& ch, S& sm, uint32_t off = 0) {
  block_signal(ch.inst, off, sm);
  if constexpr (sizeof...(Rest) > 0) chain_block_signal(ch.next, sm, off + I::kNQ);
}

// Returns true when this lane's chunk is to be enqueued by the tile's deferred enqueue (`defer`).
template <class S, class I, class... Rest>
__device__ inline bool chain_block_release(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, S& sm, uint32_t gc,
                                           uint32_t r, uint32_t chunk, uint32_t page, bool defer, uint32_t& gen,
                                           uint32_t off = 0) {
  const bool here = gc != kMarker && gc >= off && gc < off + I::kNQ;
  bool won = false;
  if (dev::ballot(here)) {
    const uint32_t qi = here ? gc - off : 0u;
    if constexpr (I::kChunks) {
      const bool w = block_release_chunks(m, ch.inst, base, here, qi, chunk, page, &sm.total[here ? gc : off], defer, gen);
      if (here) won = w;
    } else {
      Ctx<I> x{m, ch.inst, base};
      put_lanes<64>(x, &ch.inst.q[qi], here, here ? sm.base[gc] + r : 0u, make_index(chunk, page));
    }
  }
  if constexpr (sizeof...(Rest) > 0) {
    const bool w = chain_block_release(m, ch.next, base, sm, gc, r, chunk, page, defer, gen, off + I::kNQ);
    if (!here) won = w;
  }
  return won;
}

template <class S, class I, class... Rest>
__device__ inline void chain_block_enqueue_back(Chain<I, Rest...>& ch, S& sm, uint32_t off = 0) {
  if constexpr (I::kChunks) block_enqueue_back(ch.inst, off, sm);
  if constexpr (sizeof...(Rest) > 0) chain_block_enqueue_back(ch.next, sm, off + I::kNQ);
}

template <uint32_t T, class S, class I, class... Rest>
__device__ inline void chain_block_enqueue(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, const S& sm,
                                           const uint32_t (&ec)[T], const uint32_t (&er)[T], const uint32_t (&chunk)[T],
                                           const uint32_t (&gen)[T], uint32_t off = 0) {
  if constexpr (I::kChunks) {
    bool any = false;
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) any |= ec[t] != kMarker && ec[t] >= off && ec[t] < off + I::kNQ;
    if (dev::ballot(any)) block_enqueue_chunks<T>(m, ch.inst, base, off, sm, ec, er, chunk, gen);
  }
  if constexpr (sizeof...(Rest) > 0) chain_block_enqueue<T>(m, ch.next, base, sm, ec, er, chunk, gen, off + I::kNQ);
}

// enqueueInitialChunk for class qi of instance I (single lane).
template <class I>
__device__ inline void inst_preload(Memory& m, I& inst, ReuseRing& base, uint32_t qi, uint32_t chunk, int available) {
  Ctx<I> x{m, inst, base};
  ClassQueue* cq = &inst.q[qi];
  const uint32_t ppc = ppc_of(x, cq), psize = page_size_of(x, cq);
  const uint32_t avail = available < 0 ? 0u : ((uint32_t)available > ppc ? ppc : (uint32_t)available);
  const uint32_t taken = ppc - avail;
  mark_chunk_class(m, chunk, I::kAdd, psize);
  if constexpr (I::kChunks) {
    ChunkHdr* h = hdr<ChunkHdr>(m, chunk);
    const uint32_t gen = make_gen(dev::add(&m.carves, 1u), qi);
    dev::st(&h->page_size, psize);
    dev::st(&h->size, ppc);
    dev::st(count_gen(h), make_cg((int32_t)avail, gen));
    dev::st(&h->queue_pos, avail ? qp_pending(gen) : qp_not(gen));  // enqueued below when it has pages
    for (uint32_t w = 0; w < kMaskWords; ++w) {
      uint64_t bits = 0;
      for (uint32_t b = 0; b < 64; ++b) {
        const uint32_t pg = w * 64 + b;
        if (pg >= taken && pg < ppc) bits |= 1ull << b;
      }
      dev::st(&h->mask[w], bits);
    }
    dev::order();
    dev::add(&cq->carved, 1u);
    if (avail) enqueue_chunks<1>(x, cq, true, chunk, gen);  // ChunkQueue_impl.cuh:74-75
  } else {
    PageChunkHdr* h = hdr<PageChunkHdr>(m, chunk);
    dev::st(&h->page_size, psize);
    dev::st(&h->number_pages, ppc);
    dev::order();
    // the whole chunk belongs to the class: the caller's pages come back through free()
    dev::add(&cq->carved, ppc);
    const uint32_t v0 = avail ? dev::add(&cq->back, avail) : 0u;
    for (uint32_t p = taken; p < ppc; ++p) put_lanes<1>(x, cq, true, v0 + (p - taken), make_index(chunk, p));
  }
  dev::order();
  sem_signal(cq, avail);  // publish, then signal
}

// Chain routing for enqueueInitialChunk / preloadChunk: class indices run over the whole chain.
template <class I, class... Rest>
__device__ inline void chain_preload(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, uint32_t k, uint32_t chunk,
                                     int available, uint32_t ppc) {
  if (k < I::kNQ) {
    if (ppc != ((I::kChunkSize >> I::kSmallLog2) >> k) || !chunk_ok(m, chunk)) {
      set_error(m, ERR_CORRUPT);
      return;
    }
    inst_preload(m, ch.inst, base, k, chunk, available);
    return;
  }
  if constexpr (sizeof...(Rest) > 0) chain_preload(m, ch.next, base, k - I::kNQ, chunk, available, ppc);
  else set_error(m, ERR_CORRUPT);
}

template <class I, class... Rest>
__device__ inline uint32_t chain_preload_take(Memory& m, Chain<I, Rest...>& ch, ReuseRing& base, uint32_t k,
                                              int available) {
  if (k < I::kNQ) {
    uint32_t chunk = kMarker;
    if (alloc_chunks(m, ch.inst.reuse, I::kAdd, 1, chunk) != 1) return kMarker;
    inst_preload(m, ch.inst, base, k, chunk, available);
    return chunk;
  }
  if constexpr (sizeof...(Rest) > 0) return chain_preload_take(m, ch.next, base, k - I::kNQ, available);
  else {
    set_error(m, ERR_CORRUPT);
    return kMarker;
  }
}

template <class I>
inline void host_instance_geometry(InstanceGeometry& g) {
  g.kind = (uint32_t)I::kKind;
  g.num_classes = I::kNQ;
  g.queue_entries = I::kQueueEntries;
}

}  // namespace detail

template <class... Is>
struct Ouroboros {
  using MyType = Ouroboros<Is...>;
  static constexpr int kNumInstances = sizeof...(Is);
  static constexpr bool kVirtualized = ((!is_std(Is::kKind)) || ...);
  // base chunks per chunk of the first instance; queue chunks take runs of this length too
  static constexpr uint32_t kBaseAdd = std::tuple_element_t<0, std::tuple<Is...>>::kAdd;
  // Geometry tag of the chain (queue kinds, page ranges, queue capacities), kept in Memory::variant so
  // a restored copy can be checked against the pool it goes into.
  static constexpr uint32_t kChainTag = [] {
    uint32_t t = 2166136261u;
    ((t = (t ^ ((uint32_t)Is::kKind | (Is::kSmallLog2 << 4) | (Is::kNQ << 12))) * 16777619u,
      t = (t ^ Is::kQueueEntries) * 16777619u),
     ...);
    return t;
  }();

  Memory memory;
  Chain<Is...> chain;

  // ---------------------------------------------------------------- device API
  template <int W = OURO_DEFAULT_AGG>
  __device__ void* malloc(size_t size) {
    static_assert(W == 1 || W == 16 || W == 64, "aggregation width must be 1, 16 or 64");
    return detail::chain_malloc<W>(memory, chain, chain.inst.reuse, (uint64_t)size, true);
  }

  // Block-wide malloc / free: every thread of the workgroup calls together (blockDim.x a multiple of
  // 64, at most 1024), size 0 / nullptr for a thread without a request; `sm` is the block's
  // __shared__ Scratch. Requests of one class form one group for the whole block.
  static constexpr uint32_t kNumClasses = (Is::kNQ + ...);
  static constexpr uint32_t kLargestPage = [] {
    uint32_t m = 0;
    ((m = Is::kLargest > m ? Is::kLargest : m), ...);
    return m;
  }();
  static constexpr bool kAnyChunks = (Is::kChunks || ...);
  // every instance a standard page ring: block-wide takes and puts of a tile go out together
  static constexpr bool kRingPages = !kAnyChunks && !kVirtualized;
  static constexpr bool kVAPages = ((is_va(Is::kKind) && !Is::kChunks) && ...);
  static constexpr bool kAllChunks = (Is::kChunks && ...);
  static_assert(kNumClasses <= kBlockMaxClasses, "too many size classes for the block scratch");
  template <uint32_t T>
  using TileScratch = BlockScratch<kAnyChunks, T>;
  using Scratch = TileScratch<1>;

  __device__ void* malloc_block(size_t size, Scratch& sm) {
    const size_t s[1] = {size};
    void* p[1];
    malloc_tile<1>(s, p, sm);
    return p[0];
  }

  __device__ void free_block(void* ptr, Scratch& sm) {
    void* const p[1] = {ptr};
    free_tile<1>(p, sm);
  }

  // malloc_block / free_block with T requests per thread: request t of thread i is request
  // t * blockDim.x + i of a tile of T * blockDim.x, served as one block-wide group per class (one
  // semaphore claim and one front_ / back_ fetch-add per class for the whole tile), so a workgroup
  // serves T times the requests per barrier round. `sm` is a __shared__ TileScratch<T>.
  template <uint32_t T>
  __device__ void malloc_tile(const size_t (&size)[T], void* (&out)[T], TileScratch<T>& sm) {
    uint32_t gc[T], r[T];
    bool any = false;
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      gc[t] = detail::chain_alloc_class<Is...>((uint64_t)size[t]);
      if (!size[t]) gc[t] = kMarker;
      any |= gc[t] != kMarker;
    }
#if OURO_STATISTICS
    if (memory.stats_on) {
      uint32_t n = 0;
#pragma unroll
      for (uint32_t t = 0; t < T; ++t) n += dev::popc(dev::ballot(gc[t] != kMarker));
      if (dev::lane() == 0) detail::stat_add(memory, STAT_PAGE_ALLOC, n);
    }
#endif
    (void)any;
    detail::block_rank_tile<T>(sm, gc, r, kNumClasses);
    detail::chain_block_claim(memory, chain, chain.inst.reuse, sm, kNumClasses);
    __syncthreads();
    if constexpr (kRingPages) {
#pragma unroll
      for (uint32_t t = 0; t < T; ++t) out[t] = nullptr;
      detail::chain_take_ring_tile<T>(memory, chain, sm, gc, r, out);
    } else if constexpr (kVAPages) {
#pragma unroll
      for (uint32_t t = 0; t < T; ++t) out[t] = nullptr;
      uint32_t c[sizeof...(Is)][T];
      detail::chain_take_va_slots<T>(memory, chain, chain.inst.reuse, sm, gc, r, out, c);
      dev::order();  // the slots are emptied before the queue-chunk counts, whose last update may recycle one
#pragma unroll
      for (uint32_t t = 0; t < T; ++t) detail::chain_take_va_count<T>(memory, chain, chain.inst.reuse, sm, gc, r, t, c);
    } else {
#pragma unroll
      for (uint32_t t = 0; t < T; ++t)
        out[t] = detail::chain_block_take(memory, chain, chain.inst.reuse, sm, gc[t], r[t]);
    }
    // requests above the largest page: the device heap, per request
#pragma unroll
    for (uint32_t t = 0; t < T; ++t)
      if (size[t] && gc[t] == kMarker) out[t] = detail::heap_malloc(memory, (uint64_t)size[t]);
  }

  // The floor under malloc_tile, for measurement: the same class lookup, block rank in LDS and
  // barriers, and one pointer written per request, with no allocator state touched (the pointer is
  // made up from the class and rank and never dereferenced).
  template <uint32_t T>
  __device__ void null_tile(const size_t (&size)[T], void* (&out)[T], TileScratch<T>& sm) {
    uint32_t gc[T], r[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      gc[t] = detail::chain_alloc_class<Is...>((uint64_t)size[t]);
      if (!size[t]) gc[t] = kMarker;
    }
    detail::block_rank_tile<T>(sm, gc, r, kNumClasses);
    __syncthreads();
#pragma unroll
    for (uint32_t t = 0; t < T; ++t)
      out[t] = gc[t] == kMarker ? nullptr : memory.d_data + ((uint64_t)gc[t] << 27) + (uint64_t)r[t] * 16u;
  }

  // Inlined into its kernel (free_tile_in too): compiled as a call, its register save and restore went
  // through scratch (80 B per lane in the chunk kinds' persistent free) and the frees ran 10-17% slower.
  template <uint32_t T>
  __device__ __forceinline__ void free_tile(void* const (&ptr)[T], TileScratch<T>& sm) {
    free_tile_in<T>(ptr, sm);
  }

  template <uint32_t T, class S>
  __device__ __forceinline__ void free_tile_in(void* const (&ptr)[T], S& sm) {
    uint32_t chunk[T], page[T], gc[T], r[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      uint32_t page_size = 0;
      uint64_t offset = 0;
      chunk[t] = 0;
      page[t] = 0;
      gc[t] = kMarker;
      if (detail::route_free(memory, ptr[t], chunk[t], page_size, offset)) {
        gc[t] = detail::chain_free_class<Is...>(page_size, offset, page[t]);
        if (gc[t] == kMarker) detail::set_error(memory, ERR_BAD_FREE);
      }
    }
#if OURO_STATISTICS
    if (memory.stats_on) {
      uint32_t nf = 0;
#pragma unroll
      for (uint32_t t = 0; t < T; ++t) nf += dev::popc(dev::ballot(gc[t] != kMarker));
      if (dev::lane() == 0) detail::stat_add(memory, STAT_PAGE_FREE, nf);
    }
#endif
    // Each wave's stores to the blocks it frees complete at its XCD's L2 before the block barrier
    // inside block_rank_tile; after it, one agent-scope release by wave 0 writes those lines back for
    // the whole workgroup (buffer_wbl2 acts on the entire L2 of the XCD), instead of one per wave.
    dev::order();
    detail::block_rank_tile<T>(sm, gc, r, kNumClasses);
    if (threadIdx.x < kWave) {
      if (dev::ballot(dev::lane() < kNumClasses && sm.total[dev::lane()] > 0)) dev::release_block_contents();
      detail::chain_block_back(chain, sm);
    }
    __syncthreads();
    if constexpr (kRingPages) {
      detail::chain_put_ring_tile<T>(memory, chain, sm, gc, r, chunk, page);
    } else if constexpr (!kAnyChunks) {
      // VA / VL page queues: a put at a queue-chunk boundary pre-allocates the next queue chunk and
      // puts further on wait for it, so every sub-tile's pre-allocations come before any put
      detail::chain_put_pages_tile<T>(memory, chain, chain.inst.reuse, sm, gc, r, chunk, page);
    } else {
      // chunk queues without recycling defer the enqueues of the chunks the tile's frees bring back
      // to the queue, and enqueue them together: one count_ and one back_ fetch-add per class and
      // tile instead of one per wave and sub-tile (same-address atomics serialise device-wide)
      const bool defer = kAnyChunks && memory.recycle_min == kRecycleOff;
      uint32_t ec[T], gen[T];
      bool tiled = false;
      if constexpr (kAllChunks) {
        if (defer) {
          // the whole tile's bits, counts and enqueue claims, each in flight together
          bool won[T];
#pragma unroll
          for (uint32_t t = 0; t < T; ++t) {
            won[t] = false;
            gen[t] = 0;
          }
          detail::chain_release_chunks_tile<T>(memory, chain, gc, chunk, page, won, gen);
#pragma unroll
          for (uint32_t t = 0; t < T; ++t) ec[t] = won[t] ? gc[t] : kMarker;
          tiled = true;
        }
      }
      if (!tiled) {
#pragma unroll
        for (uint32_t t = 0; t < T; ++t) {
          gen[t] = 0;
          const bool won =
              detail::chain_block_release(memory, chain, chain.inst.reuse, sm, gc[t], r[t], chunk[t], page[t], defer, gen[t]);
          ec[t] = won ? gc[t] : kMarker;
        }
      }
      if constexpr (kAnyChunks) {
        if (defer) {
          // the class totals of the frees wait in sm.served for the signal; the enqueues are ranked
          // per class like requests
          if (threadIdx.x < kNumClasses) sm.served[threadIdx.x] = sm.total[threadIdx.x];
          __syncthreads();
          uint32_t er[T];
          detail::block_rank_tile<T>(sm, ec, er, kNumClasses);
          if (threadIdx.x < kWave) detail::chain_block_enqueue_back(chain, sm);
          __syncthreads();
          detail::chain_block_enqueue<T>(memory, chain, chain.inst.reuse, sm, ec, er, chunk, gen);
          if (threadIdx.x < kNumClasses) sm.total[threadIdx.x] = sm.served[threadIdx.x];
        }
      }
    }
    __syncthreads();  // every put / release of the block is done: publish, then signal
    if (threadIdx.x < kWave) detail::chain_block_signal(chain, sm);
  }

  template <int W = OURO_DEFAULT_AGG>
  __device__ void free(void* ptr) {
    static_assert(W == 1 || W == 16 || W == 64, "aggregation width must be 1, 16 or 64");
    // one lookup for every form: the dense class byte (route_free), in place of the locator walk and
    // chunk header read of Ouroboros::free (Ouroboros_impl.cuh:95-113); device-heap blocks go back to
    // the heap there, and anything else outside the chunk grid sets ERR_BAD_FREE
    uint32_t chunk = 0, page_size = 0;
    uint64_t offset = 0;
    const bool want = detail::route_free(memory, ptr, chunk, page_size, offset);
    if (dev::ballot(want)) {
      dev::release_block_contents();
      detail::chain_free<W>(memory, chain, chain.inst.reuse, page_size, want, chunk, offset);
    }
  }

  // Bulk preload of a chunk's pages into a class queue (Ouroboros.cuh:291-295, PageQueue_impl.cuh:40-53,
  // ChunkQueue_impl.cuh:72-78): pages [ppc - available_pages, ppc) of chunk `chunk_index` become free
  // pages of class `queue_index`; the first ppc - available_pages pages stay with the caller, who may
  // free them later. Single thread, before any concurrent use, like the reference. The class index
  // runs over the whole chain (the reference reaches only its first instance); pages_per_chunk must
  // be the class's. The chunk must belong to the caller (preloadChunk takes one from the pool).
  __device__ void enqueueInitialChunk(uint32_t queue_index, uint32_t chunk_index, int available_pages,
                                      uint32_t pages_per_chunk) {
    detail::chain_preload(memory, chain, chain.inst.reuse, queue_index, chunk_index, available_pages,
                          pages_per_chunk);
  }

  // Takes a chunk (run) for class `queue_index` from the pool, as a carve would, and preloads it with
  // enqueueInitialChunk. Returns the chunk index, or kMarker (ERR_OOM set) when the pool has none.
  __device__ uint32_t preloadChunk(uint32_t queue_index, int available_pages) {
    return detail::chain_preload_take(memory, chain, chain.inst.reuse, queue_index, available_pages);
  }

  // ---------------------------------------------------------------- host API
  Ouroboros() { memset(static_cast<void*>(this), 0, sizeof(*this)); }
  ~Ouroboros() { release(); }
  Ouroboros(const Ouroboros&) = delete;  // the reference's Memory double-frees on copy
  Ouroboros& operator=(const Ouroboros&) = delete;

  MyType* getDeviceMemoryManager() { return reinterpret_cast<MyType*>(memory.d_memory); }

  static PoolLayout layout(size_t instantiation_size, size_t add_beg = 0, size_t add_end = 0) {
    InstanceGeometry g[kNumInstances];
    fill_geometry<0, Is...>(g);
    return compute_layout(instantiation_size, g, kNumInstances, kStateBytes + add_beg + add_end);
  }

  // Ouroboros::initialize (MemoryInitialization.cuh:123-219). Idempotent; returns hipSuccess or the
  // first HIP error instead of exiting.
  hipError_t initialize(size_t instantiation_size, size_t additionalSizeBeginning = 0, size_t additionalSizeEnd = 0) {
    if (memory.initialized) return hipSuccess;
    static_assert(sizeof(MyType) <= kStateBytes, "manager state exceeds its 8 KiB slot");
    const PoolLayout L = layout(instantiation_size, additionalSizeBeginning, additionalSizeEnd);
    // MemoryIndex holds 23 chunk bits: a larger pool would hand out aliased page handles
    if (L.max_chunks == 0 || L.max_chunks > kMaxChunks) return hipErrorInvalidValue;
    // one allocation holds the whole state: the reference's partition of allocation_size bytes,
    // then the chunk class bytes (so a copy of snapshot_bytes() bytes is the complete allocator)
    uint8_t* d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), L.allocation_size + class_bytes(L.max_chunks) + kDirBytes);
    if (e != hipSuccess) return e;
    memory.d_memory = d;
    memory.d_data = d + L.data_offset;
    memory.allocation_size = L.allocation_size;
    memory.max_chunks = (uint32_t)L.max_chunks;
    memory.locator_bytes = L.locator_bytes;
    memory.locator = reinterpret_cast<uint32_t*>(d + L.allocation_size - L.locator_bytes);
    memory.chunk_log2 = d + L.allocation_size;
    memory.pool_bytes = L.allocation_size + class_bytes(L.max_chunks) + kDirBytes;
    memory.instantiation_size = instantiation_size;
    memory.add_beg = (uint32_t)additionalSizeBeginning;
    memory.add_end = (uint32_t)additionalSizeEnd;
    memory.stats_on = 0;
    memory.recycle_min = kRecycleOff;
    memory.carves = 0;
    // requests above the largest page go to the device heap: give it the reference's 500 MB
    // (cuda_heap_size, Parameters.h:48; MemoryInitialization.cuh:130-134) unless it is larger already
    {
      size_t heap = 0;
      if (hipDeviceGetLimit(&heap, hipLimitMallocHeapSize) == hipSuccess && heap < kHeapBytes)
        (void)hipDeviceSetLimit(hipLimitMallocHeapSize, kHeapBytes);
      (void)hipGetLastError();
    }
    e = hipMemset(memory.chunk_log2, 0, class_bytes(L.max_chunks));
    if (e != hipSuccess) {
      release();
      return e;
    }
    memory.variant = kChainTag;
    memory.next_free_chunk = 0;
    memory.qc_next = 0;
    memory.qc_add = kBaseAdd;
    // a virtualized pool keeps its top chunks for queue chunks (at most a quarter of a tiny pool)
    memory.qc_reserve = kVirtualized ? (kQCReserve < L.max_chunks / 4 ? kQCReserve : (uint32_t)(L.max_chunks / 4)) : 0u;
    memory.error = 0;
    // queue region: rings and class queues are placed downward from the locator (initNew, :42-58)
    uint8_t* end = d + L.allocation_size - L.locator_bytes;
    std::vector<uint32_t> qcs;  // initial queue chunks of VA / VL queues, in allocation order
    std::vector<std::pair<uint32_t*, uint32_t>> va_heads;  // ring entry 0 of each VA queue
    std::vector<std::pair<uint64_t*, uint64_t>> vl_heads;  // directory entry 0 of each VL queue
    uint8_t* dir = memory.chunk_log2 + class_bytes(L.max_chunks);
    place<0, Is...>(chain, end, dir, qcs, va_heads, vl_heads);
    // every queue array and ring to the empty marker, the locator to zero, the VL directories empty
    const uint64_t qbytes = (uint64_t)(d + L.allocation_size - L.locator_bytes - end);
    e = hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(end), kMarker, qbytes / 4);
    if (e == hipSuccess && kDirBytes)
      e = hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(memory.chunk_log2 + class_bytes(L.max_chunks)), kMarker,
                       kDirBytes / 4);
    if (e == hipSuccess) e = hipMemset(memory.locator, 0, L.locator_bytes);
    // initial queue chunks (PageQueueVA::init etc. take chunks 0, 1, 2, ... via allocateChunk<true>)
    if (e == hipSuccess && !qcs.empty()) {
      std::vector<uint32_t> loc((L.locator_bytes + 3) / 4, 0u);
      std::vector<uint8_t> qc(kStride);
      for (uint32_t i = 0; i < qcs.size() && e == hipSuccess; ++i) {
        const uint32_t c = qcs[i];
        memset(qc.data(), 0xFF, qc.size());
        QueueChunkHdr h{0u, 0u, 0u, kMarker, c};
        memset(qc.data(), 0, kMeta);
        memcpy(qc.data(), &h, sizeof(h));
        e = hipMemcpy(memory.d_data + (uint64_t)c * kStride, qc.data(), kStride, hipMemcpyHostToDevice);
        loc[c >> 5] |= 1u << (c & 31u);
      }
      const uint32_t used = (uint32_t)qcs.size() * memory.qc_add;
      if (e == hipSuccess) e = hipMemcpy(memory.locator, loc.data(), ((used + 31) / 32) * 4, hipMemcpyHostToDevice);
      for (size_t i = 0; i < va_heads.size() && e == hipSuccess; ++i)
        e = hipMemcpy(va_heads[i].first, &va_heads[i].second, 4, hipMemcpyHostToDevice);
      for (size_t i = 0; i < vl_heads.size() && e == hipSuccess; ++i)
        e = hipMemcpy(vl_heads[i].first, &vl_heads[i].second, 8, hipMemcpyHostToDevice);
      memory.next_free_chunk = used;
    }
    memory.initialized = 1;
    if (e == hipSuccess) e = hipMemcpy(d, this, sizeof(MyType), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) release();
    return e;
  }

  // reinitialize(overallocation_factor) (declared, never defined, in Ouroboros.cuh:83,179,283): a fresh
  // pool of overallocation_factor times the size given to initialize(), replacing this one. Every
  // block handed out before becomes invalid and the device manager moves: call it with no block live
  // and pass getDeviceMemoryManager() anew. On failure the manager is left uninitialized.
  hipError_t reinitialize(float overallocation_factor) {
    if (!memory.initialized || !(overallocation_factor > 0.0f)) return hipErrorInvalidValue;
    const size_t size = (size_t)((double)memory.instantiation_size * (double)overallocation_factor);
    const size_t beg = memory.add_beg, end = memory.add_end;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    release();
    return initialize(size, beg, end);
  }

  // Device -> host copy of the manager state (updateMemoryManagerHost, Ouroboros_impl.cuh:137-144).
  hipError_t updateHost(MyType& out) const {
    return hipMemcpy(&out, memory.d_memory, sizeof(MyType), hipMemcpyDeviceToHost);
  }

  void release() {
    if (memory.d_memory) (void)hipFree(memory.d_memory);
    memory.d_memory = nullptr;
    memory.chunk_log2 = nullptr;
    memory.initialized = 0;
  }

  // Bytes of the chunk class array behind the pool partition, and of the whole allocation (the
  // reference's pool plus that array): a copy of snapshot_bytes() from d_memory on is the complete
  // allocator state.
  static constexpr uint64_t class_bytes(uint64_t max_chunks) { return align_up(max_chunks, 256); }
  // The VL queue-chunk directories (kVLDir entries of 8 bytes per class of every linked instance)
  // follow the chunk class bytes in the same allocation.
  static constexpr uint64_t kDirBytes =
      ((is_vl(Is::kKind) ? (uint64_t)Is::kNQ * detail::kVLDir * 8ull : 0ull) + ... + 0ull);
  static constexpr size_t kHeapBytes = 500ull * 1024 * 1024;
  uint64_t snapshot_bytes() const { return memory.allocation_size + class_bytes(memory.max_chunks) + kDirBytes; }

  // Restore a copy taken with snapshot_bytes() from the pool of an identical instantiation (same
  // chain, same pool size) into this one: the copy replaces the pool, then every address inside the
  // manager state (pool, data, locator, class bytes, queue arrays and rings) is set to this pool's
  // own, since queue contents and links hold chunk indices, never addresses. `src` may be host or
  // device memory. hipErrorInvalidValue if the copy belongs to another geometry.
  hipError_t restore(const void* src) {
    if (!memory.initialized || !src) return hipErrorInvalidValue;
    std::vector<uint8_t> head(sizeof(MyType));
    hipError_t e = hipMemcpy(head.data(), src, sizeof(MyType), hipMemcpyDefault);
    if (e != hipSuccess) return e;
    MyType* h = reinterpret_cast<MyType*>(head.data());
    if (!h->memory.initialized || h->memory.allocation_size != memory.allocation_size ||
        h->memory.max_chunks != memory.max_chunks || h->memory.variant != memory.variant ||
        (uint64_t)(h->memory.d_data - h->memory.d_memory) != (uint64_t)(memory.d_data - memory.d_memory))
      return hipErrorInvalidValue;
    e = hipMemcpy(memory.d_memory, src, snapshot_bytes(), hipMemcpyDefault);
    if (e != hipSuccess) return e;
    h->memory.d_memory = memory.d_memory;
    h->memory.d_data = memory.d_data;
    h->memory.locator = memory.locator;
    h->memory.chunk_log2 = memory.chunk_log2;
    rebase<0, Is...>(h->chain, chain);
    e = hipMemcpy(memory.d_memory, head.data(), sizeof(MyType), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
  }

 private:
  // queue-array and ring addresses of `to` taken from this pool's host mirror `from`
  template <int N>
  static void rebase(Chain<>&, const Chain<>&) {}
  template <int N, class I, class... Rest>
  static void rebase(Chain<I, Rest...>& to, const Chain<I, Rest...>& from) {
    for (uint32_t i = 0; i < 16; ++i) to.inst.tab.ring[i] = from.inst.tab.ring[i];
    to.inst.reuse.ring = from.inst.reuse.ring;
    for (uint32_t i = 0; i < I::kNQ; ++i) to.inst.q[i].ring = from.inst.q[i].ring;
    rebase<N + 1, Rest...>(to.next, from.next);
  }

  template <int N>
  static void fill_geometry(InstanceGeometry*) {}
  template <int N, class I, class... Rest>
  static void fill_geometry(InstanceGeometry* g) {
    detail::host_instance_geometry<I>(g[N]);
    fill_geometry<N + 1, Rest...>(g);
  }

  template <int N>
  static void place(Chain<>&, uint8_t*&, uint8_t*&, std::vector<uint32_t>&, std::vector<std::pair<uint32_t*, uint32_t>>&,
                    std::vector<std::pair<uint64_t*, uint64_t>>&) {}
  template <int N, class I, class... Rest>
  static void place(Chain<I, Rest...>& ch, uint8_t*& end, uint8_t*& dir, std::vector<uint32_t>& qcs,
                    std::vector<std::pair<uint32_t*, uint32_t>>& va_heads,
                    std::vector<std::pair<uint64_t*, uint64_t>>& vl_heads) {
    InstanceGeometry g;
    detail::host_instance_geometry<I>(g);
    I& inst = ch.inst;
    end -= reuse_bytes();
    inst.reuse.ring = reinterpret_cast<uint32_t*>(end);
    inst.reuse.size = kReuseEntries;
    inst.reuse.count = 0;
    inst.reuse.front = inst.reuse.back = 0;
    for (uint32_t i = 0; i < I::kNQ; ++i) {
      ClassQueue& q = inst.q[i];
      const uint64_t qb = queue_bytes(g);
      end -= qb;
      q.ring = qb ? reinterpret_cast<uint32_t*>(end) : nullptr;
      inst.tab.ring[i] = q.ring;
      q.ring_size = (uint32_t)g.queue_entries;
      q.page_size = I::kSmallest << i;
      q.ppc = I::kChunkSize / q.page_size;
      q.qindex = i;
      q.sem = Sem::kInit;
      q.front = q.back = 0;
      q.count = 0;
      q.scan = 0;
      q.carved = 0;
      q.old_count = 0;
      q.front_qc = q.back_qc = q.old_qc = kMarker;
      if (!is_std(I::kKind)) {
        // queue chunks take whole runs of the first instance's chunk length
        const uint32_t c = (uint32_t)qcs.size() * kBaseAdd;
        qcs.push_back(c);
        if (is_vl(I::kKind)) {
          q.front_qc = q.back_qc = q.old_qc = c;
          // the directory's table pointer takes the unused queue-array entry of a linked queue
          uint64_t* t = reinterpret_cast<uint64_t*>(dir);
          dir += detail::kVLDir * 8ull;
          inst.tab.ring[i] = reinterpret_cast<uint32_t*>(t);
          vl_heads.emplace_back(t, ((uint64_t)0 << 32) | c);  // queue chunk 0 starts at position 0
        }
        if (is_va(I::kKind)) va_heads.emplace_back(q.ring, c);
      }
    }
    place<N + 1, Rest...>(ch.next, end, dir, qcs, va_heads, vl_heads);
  }
};

}  // namespace ouro
