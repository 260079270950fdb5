// One Ouroboros instance: NQ size classes over chunks of SMALLEST << (NQ-1) bytes, with one of the
// six queue organisations. Device side only; the host half lives in ouroboros.hpp.
//
// Every routine here is called by all active lanes of a wavefront together ("wave op"). Work that
// touches shared counters is grouped by match<G> and issued once per group by the group's lowest
// lane, so results inside one wave are assigned in lane order. G is the aggregation width:
//   W = 1  : the reference's own aggregation only. BulkSemaphore::wait groups a warp
//            (BulkSemaphore_impl.cuh:207-299) and atomicAggInc groups front_/back_
//            (Utility.cuh:76-95); those become 64-lane groups. Everything else is per lane:
//            semaphore signals, chunk-queue back_/count_, ChunkAccess claims and releases.
//   W = 16 : everything grouped inside 16-lane segments.
//   W = 64 : everything grouped across the wavefront, chunk pages claimed and released in
//            64-bit mask words, one semaphore signal per class group.
#pragma once
#include "device.hpp"
#include "layout.hpp"

namespace ouro {

// QCAP: entries of each class queue array (0 = the reference's capacity for the kind).
template <Kind K, uint32_t SMALLEST, uint32_t NQ, uint32_t QCAP = 0u>
struct Instance {
  static constexpr Kind kKind = K;
  static constexpr uint32_t kSmallest = SMALLEST;
  static constexpr uint32_t kNQ = NQ;
  static constexpr uint32_t kSmallLog2 = log2_floor(SMALLEST);
  static constexpr uint32_t kLargest = SMALLEST << (NQ - 1);       // LargestPageSize_
  static constexpr uint32_t kChunkSize = kLargest;                  // ChunkSize_ (Ouroboros.cuh:61)
  static constexpr uint32_t kAdd = kChunkSize / kBaseChunk;         // ChunkAddFactor_ (:62)
  static constexpr bool kIsBase = kAdd == 1;
  static constexpr bool kChunks = is_chunk_kind(K);
  // entries of each class's queue array (standard: slots; VA: queue-chunk pointers; VL: none)
  static constexpr uint32_t kQueueEntries =
      is_vl(K) ? 0u : (QCAP ? QCAP : (is_std(K) ? kStdEntries : kVAEntries));
  static_assert(kQueueEntries == 0 || (kQueueEntries & (kQueueEntries - 1)) == 0, "queue capacity must be a power of two");
  static_assert(kChunkSize % kBaseChunk == 0, "instance chunk must be a multiple of the base chunk");
  static_assert(kChunkSize / SMALLEST <= 512, "at most 512 pages per chunk (9 page bits)");
  static_assert(NQ <= 16, "at most 16 classes per instance");
  // the chunk class byte keeps a base chunk's offset inside its run in 3 bits
  static_assert(kAdd <= 8, "an instance chunk spans at most 8 base chunks");

  // Read-only after initialize: the class queue arrays. Kept off the lines of the class counters,
  // whose atomics slow down when plain loads of the same line keep arriving.
  struct alignas(128) Table {
    uint32_t* ring[16];
  } tab;
  ReuseRing reuse;   // d_chunk_reuse_queue
  ClassQueue q[NQ];  // d_storage_reuse_queue
};

namespace detail {

using namespace dev;

__device__ __forceinline__ void set_error(Memory& m, uint32_t bits) { or_(&m.error, bits); }

// Adds n to statistics counter i (the caller's lane issues it once for its wave or group).
__device__ __forceinline__ void stat_add(Memory& m, uint32_t i, uint64_t n) {
#if OURO_STATISTICS
  if (n && m.stats_on) add(&m.stats[i], (unsigned long long)n);
#else
  (void)m, (void)i, (void)n;
#endif
}
// Statistics for the active lanes of a wave with `want` set: one atomic by the lowest such lane.
__device__ __forceinline__ void stat_wave(Memory& m, uint32_t i, bool want) {
#if OURO_STATISTICS
  if (!m.stats_on) return;
  const uint64_t b = ballot(want);
  if (b && (int)lane() == first(b)) add(&m.stats[i], (unsigned long long)popc(b));
#else
  (void)m, (void)i, (void)want;
#endif
}

// Class byte of every base chunk of a freshly carved run (Memory::chunk_log2), published with the
// chunk's header before any of its pages.
__device__ __forceinline__ void mark_chunk_class(Memory& m, uint32_t chunk, uint32_t run, uint32_t page_size) {
  const uint32_t lg = 31u - (uint32_t)__builtin_clz(page_size);
  for (uint32_t j = 0; j < run; ++j) st(&m.chunk_log2[chunk + j], (uint8_t)(lg | (j << 5)));
}
__device__ __forceinline__ void clear_chunk_class(Memory& m, uint32_t chunk, uint32_t run) {
  for (uint32_t j = 0; j < run; ++j) st(&m.chunk_log2[chunk + j], (uint8_t)0);
}
// a bounded wait gave up; diagnostic builds also log the source line of the wait
#define OURO_TIMEOUT(mem)                                         \
  do {                                                            \
    OURO_DIAG_LOG((3ull << 60) | (unsigned long long)__LINE__);   \
    ::ouro::detail::set_error((mem), ::ouro::ERR_TIMEOUT);        \
  } while (0)

__device__ __forceinline__ uint8_t* chunk_addr(const Memory& m, uint32_t c) {
  return m.d_data + (uint64_t)c * kStride;
}
template <class T>
__device__ __forceinline__ T* hdr(const Memory& m, uint32_t c) {
  return reinterpret_cast<T*>(chunk_addr(m, c));
}
__device__ __forceinline__ uint32_t* qc_slots(const Memory& m, uint32_t c) {
  return reinterpret_cast<uint32_t*>(chunk_addr(m, c) + kMeta);
}

// A chunk index read back from device memory is only dereferenced when it names a pool chunk.
__device__ __forceinline__ bool chunk_ok(const Memory& m, uint32_t c) { return c < m.max_chunks; }

// Broadcast from the group leader of each lane (every lane of the wave must be active).
__device__ __forceinline__ int leader_of(uint64_t grp) { return first(grp); }

// ------------------------------------------------------------------ reuse ring (Queue_impl.cuh)
// IndexQueue::enqueue (Queue_impl.cuh:20-38) for every lane with want=true, positions in lane order.
// Returns whether this lane's value went in; values beyond the ring's room are refused, which sets
// ERR_REUSE_FULL unless the caller takes them back itself (strict = false). `headroom` entries of the
// ring are left free (recycled data chunks leave room for the queue chunks that share the ring).
__device__ inline bool ring_put(Memory& m, ReuseRing& r, bool want, uint32_t value, bool strict = true,
                                uint32_t headroom = 0) {
  const uint64_t mask = ballot(want);
  if (!mask) return false;
  const int l = first(mask);
  const uint32_t n = popc(mask);
  uint32_t fill = 0, pos = 0;
  if ((int)lane() == l) {
    uint32_t acc = 0;
    if (headroom) {
      // a put that keeps headroom (recycled data chunks) goes in whole or not at all, decided on a
      // plain read: the headroom absorbs puts racing past the check, and nothing is ever taken back
      // from the count, where a concurrent take could have claimed it as an entry to come
      acc = ld(&r.count) + (int32_t)n <= (int32_t)(r.size - headroom) ? n : 0u;
      if (acc) add(&r.count, (int32_t)acc);
    } else {
      const int32_t f = add(&r.count, (int32_t)n);
      const int32_t room = (int32_t)r.size - f;
      acc = room <= 0 ? 0u : (room >= (int32_t)n ? n : (uint32_t)room);
      if (acc < n) sub(&r.count, (int32_t)(n - acc));  // the ring is full: an error either way
    }
    if (acc < n && strict) set_error(m, ERR_REUSE_FULL);
    fill = acc;
    pos = acc ? add(&r.back, acc) : 0u;
    stat_add(m, STAT_CHUNK_REUSE, acc);
  }
  fill = readlane(fill, l);
  pos = readlane(pos, l);
  if (want) {
    const uint32_t rk = rank_in(mask);
    if (rk < fill) {
      uint32_t* slot = &r.ring[(pos + rk) & (r.size - 1)];
      Spin it(&m.error);
      while (cas(slot, kMarker, value) != kMarker) {
        if (it.expired()) { OURO_TIMEOUT(m); break; }
        backoff(it.it);
      }
      return true;
    }
  }
  return false;
}

// IndexQueue::dequeue (Queue_impl.cuh:54-67) for up to k entries, issued by lane `l` (uniform),
// leaving at least `keep` entries in the ring. Returns the number obtained; the j-th active lane
// (j < result) receives entry j in `out`.
__device__ inline uint32_t ring_take(Memory& m, ReuseRing& r, uint32_t k, int l, uint32_t me, uint32_t& out,
                                     uint32_t keep = 0) {
  uint32_t t = 0, pos = 0;
  if ((int)lane() == l) {
    const int32_t old = sub(&r.count, (int32_t)k) - (int32_t)keep;
    t = old <= 0 ? 0u : ((uint32_t)old >= k ? k : (uint32_t)old);
    if (t < k) add(&r.count, (int32_t)(k - t));
    pos = t ? add(&r.front, t) : 0u;
  }
  t = readlane(t, l);
  pos = readlane(pos, l);
  if (me < t) {
    uint32_t* slot = &r.ring[(pos + me) & (r.size - 1)];
    Spin it(&m.error);
    uint32_t v;
    while ((v = xchg(slot, kMarker)) == kMarker) {
      if (it.expired()) { OURO_TIMEOUT(m); break; }
      backoff(it.it);
    }
    out = v;
  }
  return t;
}

// allocateChunk<QUEUECHUNK> (Ouroboros.cuh:108-136) for k chunks at once: the reuse ring first
// (FIFO), then the bump pointer, exactly the order k sequential calls would produce. The j-th active
// lane (j < k) receives chunk j. Validity follows the reference's predicate chunk + add < maxChunks
// (:133) below the queue-chunk reserve (the top m.qc_reserve chunks of a virtualized pool); queue
// chunks (`queue`) that find the bump region spent take from the reserve, so a full pool can still
// advance its queues. Unlike the reference, an invalid chunk is never handed out. Returns the valid
// prefix length.
__device__ inline uint32_t alloc_chunks(Memory& m, ReuseRing& r, uint32_t add_factor, uint32_t k, uint32_t& chunk,
                                        bool queue = false) {
  const uint64_t act = active();
  const uint32_t me = rank_in(act);
  const int l = first(act);
  uint32_t got = 0;
  // data carves leave the last few recycled chunks of a virtualized pool to its queue chunks
  const uint32_t t = ring_take(m, r, k, l, me, got, queue ? 0u : m.qc_reserve / 2);
  uint32_t base = 0;
  if ((int)lane() == l && k > t) base = add(&m.next_free_chunk, (k - t) * add_factor);
  base = readlane(base, l);
  const uint64_t limit = (uint64_t)m.max_chunks - m.qc_reserve;
  bool valid = false;
  if (me < t) {
    chunk = got;
    valid = chunk_ok(m, got);
  } else if (me < k) {
    chunk = base + (me - t) * add_factor;
    valid = (uint64_t)chunk + add_factor < limit;
  }
  if (queue) {
    // bump lanes that came up empty continue in the reserve, in rank order
    const uint64_t miss = ballot(me < k && !valid);
    if (miss) {
      uint32_t rbase = 0;
      if ((int)lane() == l) rbase = add(&m.qc_next, popc(miss));
      rbase = readlane(rbase, l);
      if (me < k && !valid) {
        const uint32_t j = rbase + rank_in(miss);
        chunk = (uint32_t)limit + j * add_factor;
        valid = (uint64_t)chunk + add_factor <= (uint64_t)m.max_chunks;
      }
    }
  }
  if (me >= t && me < k && valid) or_(&m.locator[chunk >> 5], 1u << (chunk & 31u));  // ChunkLocator::initChunkIndex
  const uint64_t okm = ballot(me < k && valid);
  // valid chunks form a prefix in active-rank order
  const uint32_t k_ok = popc(okm);
  if (k_ok < k && (int)lane() == l) set_error(m, ERR_OOM);
  if ((int)lane() == l) stat_add(m, STAT_CHUNK_ALLOC, k_ok);
  return k_ok;
}

// Lane id of the j-th active lane.
__device__ __forceinline__ int lane_of_rank(uint64_t act, uint32_t j) { return (int)nth_set_bit(act, j); }


// ------------------------------------------------------------------ queue chunks (VA / VL)
// QueueChunk::initializeChunk (QueueChunk.cuh:221-225) plus a wave-cooperative clean of its 2048
// slots (d_cleanChunks, MemoryInitialization.cuh:21-38, does this for the whole pool at start-up).
__device__ inline void qc_init_lanes(Memory& m, bool mine, uint32_t qc, uint32_t vstart) {
  const uint64_t act = active();
  const uint32_t me = rank_in(act), nact = popc(act);
  uint64_t todo = ballot(mine);
  if (mine) OURO_DIAG_ADD(11, 1);
  while (todo) {
    const int l = first(todo);
    todo &= todo - 1;
    const uint32_t c = readlane(qc, l);
    uint32_t* s = qc_slots(m, c);
    for (uint32_t i = me; i < kQCSpots; i += nact) st(&s[i], kMarker);
  }
  if (mine) {
    QueueChunkHdr* h = hdr<QueueChunkHdr>(m, qc);
    st(&h->page_size, 0u);
    st(&h->count, 0u);
    st(&h->virtual_start, vstart);
    st(&h->next, kMarker);
    st(&h->chunk_index, qc);
  }
  order();
}

// Empty <=> A == 2048 and B == 0 (QueueChunk.cuh:185-203).
__device__ __forceinline__ bool qc_empty_after_enqueue(uint32_t old, uint32_t added) {
  return old + added == kQCSpots;
}
__device__ __forceinline__ bool qc_empty_after_dequeue(uint32_t old) { return old - (1u << 16) == kQCSpots; }

template <class I>
struct Ctx {
  Memory& m;
  I& inst;
  ReuseRing& base_ring;  // d_base_chunk_reuse_queue: queue chunks always come from the base ring
};

// Class geometry from the queue's index in its instance (compile-time sizes, no memory access).
template <class I>
__device__ __forceinline__ uint32_t class_index(Ctx<I>& x, const ClassQueue* cq) {
  return (uint32_t)(cq - x.inst.q);
}
template <class I>
__device__ __forceinline__ uint32_t ppc_of(Ctx<I>& x, const ClassQueue* cq) {
  return (I::kChunkSize >> I::kSmallLog2) >> class_index(x, cq);
}
template <class I>
__device__ __forceinline__ uint32_t page_size_of(Ctx<I>& x, const ClassQueue* cq) {
  return I::kSmallest << class_index(x, cq);
}
template <class I>
__device__ __forceinline__ uint32_t* ring_of(Ctx<I>& x, const ClassQueue* cq) {
  return x.inst.tab.ring[class_index(x, cq)];
}

// ---- VL queue-chunk directory
// A linked queue finds the queue chunk of a position by walking `next` links from back_ptr_ or
// front_ptr_, and the link to queue chunk j + 1 appears only once the put at chunk j's slot 0 has
// itself walked to chunk j: with a million puts in flight, a class's ~50 new links per kernel form
// one chain and each put walks tens of links. The directory caches, per class, the (virtual start,
// queue chunk) pair of the kVLDir most recent queue chunks, entry (virtual start / 2048) mod kVLDir,
// stored when the queue chunk has been initialised (before it is linked). While a put, take or
// removal of position v is pending, the queue chunk holding v cannot be reclaimed (fewer than 2048
// of its positions are dequeued), so an entry whose virtual start equals v's names exactly that
// queue chunk. Any other entry (not published yet, or overwritten by a queue chunk kVLDir * 2048
// positions later) sends the caller to the walk: the links stay the queue's truth. Scans, which may
// read positions already dequeued, check the queue chunk's own virtual start as well.
constexpr uint32_t kVLDir = 4096;
constexpr uint64_t kVLDirEmpty = ~0ull;
__device__ __forceinline__ uint64_t vl_dir_entry(uint32_t vstart, uint32_t qc) { return ((uint64_t)vstart << 32) | qc; }
template <class I>
__device__ __forceinline__ uint64_t* vl_dir_slot(Ctx<I>& x, const ClassQueue* cq, uint32_t vstart) {
  return reinterpret_cast<uint64_t*>(x.inst.tab.ring[class_index(x, cq)]) + ((vstart / kQCSpots) & (kVLDir - 1));
}
// The queue chunk with virtual start vb from the directory, else kMarker. With `wait`, an entry that
// still holds an older queue chunk (vb's is not published yet) is waited for.
template <class I>
__device__ inline uint32_t vl_dir_find(Ctx<I>& x, const ClassQueue* cq, uint32_t vb, bool wait) {
  const uint64_t* e = vl_dir_slot(x, cq, vb);
  Spin it(&x.m.error);
  while (true) {
    const uint64_t w = ld(e);
    const uint32_t vs = (uint32_t)(w >> 32);
    if (w != kVLDirEmpty && vs == vb) return (uint32_t)w;
    // overwritten by a later queue chunk: only the walk can find vb's
    if (!wait || (w != kVLDirEmpty && vs - vb < 0x80000000u)) {
      OURO_DIAG_ADD(1, 1);
      return kMarker;
    }
    OURO_DIAG_ADD(2, 1);
    if (it.expired()) return kMarker;  // the walk waits on the links with its own bound
    backoff(it.it);
  }
}

// ---- VL pointer maintenance (QueueChunk_impl.cuh:440-551)
template <class I>
__device__ inline void vl_set_back(Ctx<I>& x, ClassQueue* cq, uint32_t c) {
  Spin it(&x.m.error);
  while (true) {
    QueueChunkHdr* h = hdr<QueueChunkHdr>(x.m, c);
    const uint32_t nx = ld(&h->next);
    if (!chunk_ok(x.m, nx)) break;
    if (cas(&cq->back_qc, c, nx) != c) break;
    if ((ld(&hdr<QueueChunkHdr>(x.m, nx)->count) & 0xFFFFu) == kQCSpots) c = nx;
    else break;
    if (it.expired()) { OURO_TIMEOUT(x.m); break; }
  }
}
// front_ptr_ moves from an emptied queue chunk c over every queue chunk after it that is emptied too
// (all 2048 positions dequeued), one CAS per move of up to kVLAhead queue chunks: the queue chunks
// after c and their counts come from the directory, kVLAhead loads in flight at a time, so the lane
// that empties the oldest queue chunk of a long emptied stretch does not walk it link by link. A queue
// chunk before the new front_ptr_ has all its positions dequeued, which is final; the queue chunk the
// move stops at is checked again after the move (its last dequeue may have seen the old front_ptr_ and
// left the move to us). Returns the number of queue chunks passed.
constexpr uint32_t kVLAhead = 8;
template <class I>
__device__ inline uint32_t vl_set_front(Ctx<I>& x, ClassQueue* cq, uint32_t c) {
  uint32_t moved = 0;
  uint32_t vs = ld(&hdr<QueueChunkHdr>(x.m, c)->virtual_start);
  while (true) {
    uint32_t e[kVLAhead], cnt[kVLAhead];
#pragma unroll
    for (uint32_t j = 0; j < kVLAhead; ++j) {
      const uint32_t want = vs + (j + 1u) * kQCSpots;
      const uint64_t w = ld(vl_dir_slot(x, cq, want));
      e[j] = (w != kVLDirEmpty && (uint32_t)(w >> 32) == want && chunk_ok(x.m, (uint32_t)w)) ? (uint32_t)w : kMarker;
    }
    uint32_t p = 0;  // directory entries found, in a row
#pragma unroll
    for (uint32_t j = 0; j < kVLAhead; ++j)
      if (p == j && e[j] != kMarker) ++p;
#pragma unroll
    for (uint32_t j = 0; j < kVLAhead; ++j) cnt[j] = j < p ? ld(&hdr<QueueChunkHdr>(x.m, e[j])->count) : 0u;
    uint32_t k = 0;  // of those, emptied, in a row
#pragma unroll
    for (uint32_t j = 0; j < kVLAhead; ++j)
      if (k == j && j < p && cnt[j] == kQCSpots) ++k;
    uint32_t nx, step;
    bool final = false;
    if (p == 0) {  // not in the directory: follow the link
      nx = ld(&hdr<QueueChunkHdr>(x.m, c)->next);
      if (!chunk_ok(x.m, nx)) break;
      step = 1;
    } else if (k == p) {
      nx = e[p - 1];
      step = p;
      final = true;
    } else {
      nx = e[k];
      step = k + 1;
    }
    if (cas(&cq->front_qc, c, nx) != c) break;
    moved += step;
    if (!final && ld(&hdr<QueueChunkHdr>(x.m, nx)->count) != kQCSpots) break;
    c = nx;
    vs += step * kQCSpots;
  }
  return moved;
}
// setOldPointer: hold back the LARGEST_OLD_COUNT_VALUE most recently emptied queue chunks, recycle
// the rest through the base reuse ring. Requesting lanes go one at a time in lane order, each
// recycling all of its queue chunks in list order, so the base ring receives them in an order a
// sequential replay reproduces. The active lanes recycle together: the lane of rank j finds the j-th
// held queue chunk in the directory, one CAS moves old_ptr_ over the batch, and one ring put (lane
// order = list order) returns it; a held queue chunk missing from the directory is taken by the link
// walk, one at a time.
template <class I>
__device__ inline void vl_set_old(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t free_count) {
  if (want && free_count) {
    uint32_t cur_old = ld(&cq->old_count);
    if (cur_old < kOldHold) {
      cur_old = add(&cq->old_count, free_count);
      if (cur_old < kOldHold) {
        if (cur_old + free_count > kOldHold) {
          free_count = cur_old + free_count - kOldHold;
          sub(&cq->old_count, free_count);
        } else {
          free_count = 0;
        }
      }
    }
  } else {
    free_count = 0;
  }
  const uint64_t act = active();
  const uint32_t me = rank_in(act), nact = popc(act);
  uint64_t todo = ballot(free_count > 0);
  while (todo) {
    const int l = first(todo);
    todo &= todo - 1;
    ClassQueue* lq = reinterpret_cast<ClassQueue*>(readlane64(reinterpret_cast<uint64_t>(cq), l));
    uint32_t left = readlane(free_count, l);
    Spin it(&x.m.error);
    while (left > 0) {
      uint32_t cur = 0, vs = 0;
      if ((int)lane() == l) {
        cur = ld(&lq->old_qc);
        vs = chunk_ok(x.m, cur) ? ld(&hdr<QueueChunkHdr>(x.m, cur)->virtual_start) : 0u;
      }
      cur = readlane(cur, l);
      vs = readlane(vs, l);
      // ranks 1..b look up queue chunks 1..b of the held list (chunk b becomes old_ptr_)
      const uint32_t b = chunk_ok(x.m, cur) ? (left < nact - 1u ? left : nact - 1u) : 0u;
      uint32_t mine = me == 0 ? cur : kMarker;
      bool miss = false;
      if (me >= 1 && me <= b) {
        const uint32_t wvs = vs + me * kQCSpots;
        const uint64_t w = ld(vl_dir_slot(x, lq, wvs));
        mine = (w != kVLDirEmpty && (uint32_t)(w >> 32) == wvs && chunk_ok(x.m, (uint32_t)w)) ? (uint32_t)w : kMarker;
        miss = mine == kMarker;
      }
      uint32_t moved = 0;
      if (b > 0 && !ballot(miss)) {
        const uint32_t nx = readlane(mine, lane_of_rank(act, b));
        if ((int)lane() == l) moved = cas(&lq->old_qc, cur, nx) == cur ? b : 0u;
        moved = readlane(moved, l);
        ring_put(x.m, x.base_ring, moved && me < b, mine);
      } else {
        // one queue chunk by its link
        if ((int)lane() == l) {
          const uint32_t nx = chunk_ok(x.m, cur) ? ld(&hdr<QueueChunkHdr>(x.m, cur)->next) : kMarker;
          moved = chunk_ok(x.m, nx) && cas(&lq->old_qc, cur, nx) == cur ? 1u : 0u;
        }
        moved = readlane(moved, l);
        ring_put(x.m, x.base_ring, moved && (int)lane() == l, cur);
      }
      left -= moved;
      uint32_t gone = 0;
      if ((int)lane() == l && left > 0) gone = it.expired();
      if (readlane(gone, l)) {
        OURO_TIMEOUT(x.m);
        break;
      }
    }
  }
}

// Walk the linked queue chunks from `start` to the one holding virtual position v
// (QueueChunk::locateQueueChunkForPosition, QueueChunk_impl.cuh:386-414).
template <class I>
__device__ inline uint32_t vl_locate(Ctx<I>& x, uint32_t start, uint32_t v) {
  uint32_t c = start;
  Spin it(&x.m.error);
  if (!chunk_ok(x.m, c)) {
    set_error(x.m, ERR_VSTART);
    return kMarker;
  }
  const uint32_t vb = v - v % kQCSpots;  // virtual start of the queue chunk holding v
  while (true) {
    // virtual starts are compared as full 32-bit positions, whose differences wrap consistently
    // (queue-chunk numbers, v / 2048, only span 21 bits)
    const uint32_t vs = ld(&hdr<QueueChunkHdr>(x.m, c)->virtual_start);
    if (vs == vb) break;
    if (vs - vb < 0x80000000u) return kMarker;  // already past v: its queue chunk is gone
    uint32_t nx;
    while ((nx = ld(&hdr<QueueChunkHdr>(x.m, c)->next)) == kMarker) {
      if (it.expired()) { OURO_TIMEOUT(x.m); return kMarker; }
      backoff(it.it);
    }
    if (!chunk_ok(x.m, nx)) {
      set_error(x.m, ERR_VSTART);
      return kMarker;
    }
    c = nx;
    OURO_DIAG_ADD(10, 1);
    if (it.expired()) { OURO_TIMEOUT(x.m); return kMarker; }
  }
  return c;
}

// VA pointer-ring entry left by a reclaimed queue chunk: every position of virtual queue chunk vqc
// was enqueued and dequeued. Chunk indices stay below 2^23 and vqc below 2^21, so a tag is never
// mistaken for an index or a marker. (The reference writes DeletionMarker here and cannot tell a
// consumed queue chunk from one not yet published.)
__device__ __forceinline__ uint32_t qc_tag(uint32_t vqc) { return 0x80000000u | (vqc & 0x1FFFFFu); }
__device__ __forceinline__ bool is_qc_tag(uint32_t e) { return (e >> 21) == 0x400u; }

// VA: queue chunk that holds virtual position v (accessQueueElement, PageQueueVA_impl.cuh:231-248).
template <class I>
__device__ inline uint32_t va_wait_qc(Ctx<I>& x, ClassQueue* cq, uint32_t v) {
  const uint32_t id = (v / kQCSpots) & (I::kQueueEntries - 1);
  uint32_t qc;
  Spin it(&x.m.error);
  while ((qc = ld(&ring_of(x, cq)[id])) == kMarker || is_qc_tag(qc)) {
    if (it.expired()) { OURO_TIMEOUT(x.m); return kMarker; }
    backoff(it.it);
  }
  if (!chunk_ok(x.m, qc) || ld(&hdr<QueueChunkHdr>(x.m, qc)->virtual_start) / kQCSpots != v / kQCSpots) {
    set_error(x.m, ERR_VSTART);
    return kMarker;
  }
  return qc;
}

// ------------------------------------------------------------------ chunk-queue front
// Chunk queues remove exhausted chunks wherever they sit (a tombstone in the slot); front_ then
// trails the oldest live entry. compact() moves it over dead positions: tombstones, a VA queue
// chunk whose pointer-ring entry holds its reclaimed tag, and for VL every queue chunk that
// front_ptr_ has already left (those are empty, all 2048 positions dequeued). Any lane may help;
// each call does bounded work and stops at the first position it cannot prove dead.
constexpr uint32_t kCompactSteps = 256;
// Slots a scan or a compaction step loads at once.
constexpr uint32_t kPeek = 8;

template <class I>
__device__ inline void compact(Ctx<I>& x, ClassQueue* cq) {
  constexpr Kind K = I::kKind;
  static_assert(is_chunk_kind(K), "page queues dequeue strictly at front_");
  uint32_t f = ld(&cq->front), lost = 0;
  const uint32_t b = ld(&cq->back);
  for (uint32_t i = 0; i < kCompactSteps && f != b && b - f < 0x80000000u; ++i) {
    uint32_t nf = f + 1u;
    if constexpr (is_std(K)) {
      // up to kPeek positions per CAS
      uint32_t e[kPeek];
      const uint32_t n = b - f < kPeek ? b - f : kPeek;
#pragma unroll
      for (uint32_t j = 0; j < kPeek; ++j) e[j] = j < n ? ld(&ring_of(x, cq)[(f + j) & (I::kQueueEntries - 1)]) : kMarker;
      uint32_t k = 0;
#pragma unroll
      for (uint32_t j = 0; j < kPeek; ++j)
        if (k == j && e[j] == kTomb) ++k;
      if (k == 0) return;
      nf = f + k;
    } else if constexpr (is_va(K)) {
      const uint32_t e = ld(&ring_of(x, cq)[(f / kQCSpots) & (I::kQueueEntries - 1)]);
      if (e == qc_tag(f / kQCSpots)) {
        nf = (f / kQCSpots + 1u) * kQCSpots;
      } else if (!chunk_ok(x.m, e) || ld(&hdr<QueueChunkHdr>(x.m, e)->virtual_start) / kQCSpots != f / kQCSpots ||
                 ld(&qc_slots(x.m, e)[f % kQCSpots]) != kTomb) {
        return;
      }
    } else {
      const uint32_t c = ld(&cq->front_qc);
      if (!chunk_ok(x.m, c)) return;
      // full 32-bit virtual starts: their differences wrap consistently with the positions
      const uint32_t vs = ld(&hdr<QueueChunkHdr>(x.m, c)->virtual_start), fb = f - f % kQCSpots;
      if (vs != fb) {
        if (vs - fb >= 0x80000000u) return;  // front_ptr_ still behind f: its move is in flight
        nf = vs;
      } else if (ld(&qc_slots(x.m, c)[f % kQCSpots]) != kTomb) {
        OURO_DIAG_ADD(14, 1);
        return;
      }
    }
    if (const uint32_t seen = cas(&cq->front, f, nf); seen != f) {
      OURO_DIAG_ADD(15, 1);
      // another lane moved front_: carry on from there (a few times at most), unless it went back
      if (seen - f >= 0x80000000u || ++lost > 8) return;
      f = seen;
      continue;
    }
    // the ring slots become writable again for positions + ring_size (only the winner of the move
    // writes them; the next writer of a slot waits for the marker with a CAS)
    if constexpr (is_std(K)) {
      for (uint32_t j = f; j != nf; ++j) st(&ring_of(x, cq)[j & (I::kQueueEntries - 1)], kMarker);
      if (nf - f < kPeek && nf != b) return;  // stopped at a position it could not prove dead
    }
    f = nf;
  }
}

// compact() for the active lanes of a wave together (all of them must call): each round looks at up
// to 16 positions per active lane from front_ and moves front_ over the leading dead ones with one CAS
// (a reclaimed VA queue chunk or a VL queue chunk front_ptr_ has left is a single move). It stops at
// the first position it cannot prove dead or after kWaveCompactSteps moves (then it returns true).
// Every remover of a class
// compacts, so the front_ CAS is contended and its rate bounds how fast front_ advances: each
// successful CAS should pass as many positions as it can. A wave whose CAS lost to a compactor that
// already passed every position this wave saw dead leaves the rest to that winner; otherwise it
// retries from the new front_ (a few times at most).
constexpr uint32_t kWaveCompactSteps = 8192;

template <uint32_t kPer = 16, class I>  // kPer: positions per lane and round
__device__ inline bool compact_wave_pass(Ctx<I>& x, ClassQueue* cq) {
  constexpr Kind K = I::kKind;
  static_assert(is_chunk_kind(K), "page queues dequeue strictly at front_");
  const uint64_t act = active();
  const int L = first(act);
  const uint32_t nact = popc(act), me = rank_in(act);
  const bool lead = (int)lane() == L;
  uint32_t moves = 0, lost = 0;
  while (moves < kWaveCompactSteps && lost < 8) {
    uint32_t f = 0, n = 0, jump = kMarker;
    const uint32_t* sl = nullptr;
    if (lead) {
      f = ld(&cq->front);
      n = ld(&cq->back) - f;
      if (n >= 0x80000000u) n = 0;
      if (n > nact * kPer) n = nact * kPer;
      if (n > kWaveCompactSteps - moves) n = kWaveCompactSteps - moves;
      if constexpr (is_va(K)) {
        const uint32_t e = ld(&ring_of(x, cq)[(f / kQCSpots) & (I::kQueueEntries - 1)]);
        if (e == qc_tag(f / kQCSpots)) jump = (f / kQCSpots + 1u) * kQCSpots;
        else if (chunk_ok(x.m, e) && ld(&hdr<QueueChunkHdr>(x.m, e)->virtual_start) / kQCSpots == f / kQCSpots)
          sl = qc_slots(x.m, e) + f % kQCSpots;
        else n = 0;
      } else if constexpr (is_vl(K)) {
        const uint32_t c = ld(&cq->front_qc);
        if (!chunk_ok(x.m, c)) {
          n = 0;
        } else {
          const uint32_t vs = ld(&hdr<QueueChunkHdr>(x.m, c)->virtual_start), fb = f - f % kQCSpots;
          if (vs == fb) sl = qc_slots(x.m, c) + f % kQCSpots;
          else if (vs - fb >= 0x80000000u) n = 0;  // front_ptr_ still behind f: its move is in flight
          else jump = vs;
        }
      }
      if constexpr (!is_std(K)) {
        if (n > kQCSpots - f % kQCSpots) n = kQCSpots - f % kQCSpots;
      }
    }
    f = readlane(f, L);
    n = readlane(n, L);
    jump = readlane(jump, L);
    sl = reinterpret_cast<const uint32_t*>(readlane64(reinterpret_cast<uint64_t>(sl), L));
    if (n == 0) break;
    uint32_t nf = jump;
    if (jump == kMarker) {
      // position r * nact + rank for the lane of that rank
      uint32_t k = 0, e[kPer];
#pragma unroll
      for (uint32_t r = 0; r < kPer; ++r) {  // all loads first: one round trip per round
        const uint32_t i = r * nact + me;
        e[r] = kMarker;
        if (i < n) e[r] = is_std(K) ? ld(&ring_of(x, cq)[(f + i) & (I::kQueueEntries - 1)]) : ld(&sl[i]);
      }
#pragma unroll
      for (uint32_t r = 0; r < kPer; ++r) {
        // leading dead ranks of this round: active lanes below the first live one
        const uint64_t alive = act & ~ballot(e[r] == kTomb);
        const uint32_t run = alive ? popc(act & ((1ull << __builtin_ctzll(alive)) - 1ull)) : nact;
        if (k == r * nact) k += run;  // still inside the leading dead stretch
      }
      if (k > n) k = n;
      if (k == 0) break;
      nf = f + k;
    }
    uint32_t seen = 0;
    if (lead) {
      OURO_DIAG_ADD(14, 1);
      seen = cas(&cq->front, f, nf);
      if (seen != f) OURO_DIAG_ADD(15, 1);
    }
    seen = readlane(seen, L);
    if (seen != f) {
      // another compactor moved front_. If it passed every position this wave saw dead (and this
      // wave stopped at one it could not), the winner carries on from there; else retry from seen.
      if (seen - f >= 0x80000000u) break;
      if (jump == kMarker && nf - f < n && seen - f >= nf - f) break;
      ++lost;
      continue;
    }
    // the ring slots become writable again for positions + ring_size; only the winner of the move
    // writes them, and the next writer of a slot waits for the marker with a CAS
    if constexpr (is_std(K)) {
#pragma unroll
      for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = r * nact + me;
        if (i < nf - f) st(&ring_of(x, cq)[(f + i) & (I::kQueueEntries - 1)], kMarker);
      }
    }
    if (jump != kMarker) {
      moves += 1;
    } else {
      moves += nf - f;
      if (nf - f < n) break;  // stopped at a live position
    }
  }
  return moves >= kWaveCompactSteps;  // stopped at the step limit, not at a live position
}

// ------------------------------------------------------------------ slot sequences
// ---- standard page-ring slots
// A page ring's slot has one writer and one reader per lap: the put and the take of its position,
// each unique (back_ / front_ fetch-adds). The slot itself is the hand-over flag: the put waits for
// MARKER (the previous lap's take is done) and stores the page; the take waits for a page and stores
// MARKER. Between a side's load and its store nobody else may write the slot, so plain device-coherent
// loads and stores suffice, and the lanes of a class, whose positions are consecutive, hit
// consecutive slots: a wave's accesses coalesce where per-lane atomics would serialise per address.
// (The reference exchanges / CASes each slot, PageQueue_impl.cuh:22-36, 84-104.)
__device__ inline void ring_slot_put(Memory& m, uint32_t* slot, uint32_t val) {
  Spin it(&m.error);
  while (ld(slot) != kMarker) {
    if (it.expired()) { OURO_TIMEOUT(m); return; }
    backoff(it.it);
  }
  st(slot, val);
}
__device__ inline uint32_t ring_slot_take(Memory& m, uint32_t* slot) {
  uint32_t val;
  Spin it(&m.error);
  while ((val = ld(slot)) == kMarker) {
    if (it.expired()) { OURO_TIMEOUT(m); return kMarker; }
    backoff(it.it);
  }
  st(slot, kMarker);
  return val;
}

// put_lanes: every lane with want=true stores `val` at virtual position v of its queue cq.
//   PQ      : ring slot handed over with plain coherent loads and stores (below; VA / VL page
//             queues hand their queue-chunk slots over the same way)
//   CQ      : ring slot CAS (PageQueue_impl.cuh:22-36, ChunkQueue_impl.cuh:20-38)
//   VAPQ/VACQ: queue chunk via the pointer ring; position 0 of a queue chunk pre-allocates the next
//              one (PageQueueVA_impl.cuh:196-229, ChunkQueueVA_impl.cuh:276-306)
//   VLPQ/VLCQ: linked queue chunks (QueueChunk_impl.cuh:137-185), with the lockstep traversal of
//              guaranteeWarpSyncPerChunk (:7-72) so a lane never waits on a link that a lane of
//              the same wave sets later.
// Queue-chunk allocation happens in a pre-phase in lane order, so chunk indices are handed out
// in the same order a sequential replay would. A queue chunk's counter (QueueChunk count_) takes one
// atomic per queue chunk of a G-lane group (G = 1: one per lane, as the reference), led by the
// group's lowest lane.
//
// put_prealloc is the pre-phase alone (it waits for nothing); put_lanes_after the rest, taking the
// queue chunk the pre-phase allocated (qc, VL links it). A caller with several positions per lane
// runs every pre-phase before any put, so that no put waits on a pre-allocation its own wave has
// still to make.
template <class I>
__device__ inline uint32_t put_prealloc(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t v) {
  constexpr Kind K = I::kKind;
  uint32_t qc = kMarker;
  if constexpr (!is_std(K)) {
    // position 0 of a queue chunk allocates the next queue chunk
    const bool need = want && (v % kQCSpots) == 0;
    if (ballot(need)) {
      uint32_t got = kMarker;
      // lanes that need one form a sub-wave; alloc_chunks hands them out in lane order
      if (need) {
        const uint32_t k = popc(ballot(true));
        // queue chunks share the base ring with the first instance's chunks: same run length
        const uint32_t ok = alloc_chunks(x.m, x.base_ring, x.m.qc_add, k, got, true);
        (void)ok;
        qc = got;
        if (rank_in(ballot(true)) >= ok) qc = kMarker;
      }
      const bool have = need && qc != kMarker;
      qc_init_lanes(x.m, have, qc, v + kQCSpots);
      // the initialised queue chunk goes into the directory before it is linked
      if constexpr (is_vl(K)) {
        if (have) st(vl_dir_slot(x, cq, v + kQCSpots), vl_dir_entry(v + kQCSpots, qc));
      }
      if constexpr (is_va(K)) {
        if (have) {
          // the entry still names the queue chunk ring_size positions back: publish only over a
          // marker or a reclaimed tag, never over a queue chunk with live entries
          uint32_t* e = &ring_of(x, cq)[((v / kQCSpots) + 1) & (I::kQueueEntries - 1)];
          uint32_t cur = ld(e);
          bool placed = false;
          while (!placed && (cur == kMarker || is_qc_tag(cur))) {
            const uint32_t seen = cas(e, cur, qc);
            placed = seen == cur;
            cur = seen;
          }
          if (!placed) set_error(x.m, ERR_QUEUE_FULL);
        }
      }
    }
  }
  return qc;
}

template <int G, class I>
__device__ inline void put_lanes_after(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t v, uint32_t val, uint32_t qc);

template <int G, class I>
__device__ inline void put_lanes(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t v, uint32_t val) {
  const uint32_t qc = put_prealloc(x, cq, want, v);
  put_lanes_after<G>(x, cq, want, v, val, qc);
}

template <int G, class I>
__device__ inline void put_lanes_after(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t v, uint32_t val, uint32_t qc) {
  constexpr Kind K = I::kKind;
  if constexpr (is_std(K) && !is_chunk_kind(K)) {
    if (want) ring_slot_put(x.m, &ring_of(x, cq)[v & (I::kQueueEntries - 1)], val);
    order();  // the slots are written before the caller signals the pages
  } else if constexpr (is_std(K)) {
    if (want) {
      uint32_t* slot = &ring_of(x, cq)[v & (I::kQueueEntries - 1)];
      Spin it(&x.m.error);
      while (cas(slot, kMarker, val) != kMarker) {
        // a chunk queue's slot may still hold the tombstone of position v - ring_size
        if constexpr (is_chunk_kind(K)) compact(x, cq);
        if (it.expired()) { OURO_TIMEOUT(x.m); break; }
        backoff(it.it);
      }
    }
  } else {
    if constexpr (is_va(K)) {
      bool reclaim = false;
      uint32_t victim = 0;
      uint32_t c = kMarker;
      if (want) {
        c = va_wait_qc(x, cq, v);
        if (c != kMarker) {
          uint32_t* slot = &qc_slots(x.m, c)[v % kQCSpots];
          if constexpr (is_chunk_kind(K)) {
            Spin it(&x.m.error);
            while (cas(slot, kMarker, val) != kMarker) {
              if (it.expired()) { OURO_TIMEOUT(x.m); break; }
              backoff(it.it);
            }
          } else {
            ring_slot_put(x.m, slot, val);  // one writer and one reader per position, as a page ring
          }
        }
      }
      // the slots are written before the queue-chunk count, whose last update may recycle it
      if constexpr (!is_chunk_kind(K)) order();
      const uint64_t grp = match<G>(c != kMarker ? (uint64_t)c : ~0ull);
      if (c != kMarker && (int)lane() == first(grp)) {
        const uint32_t added = popc(grp) * 0x10001u;
        const uint32_t old = add(&hdr<QueueChunkHdr>(x.m, c)->count, added);
        if (qc_empty_after_enqueue(old, added)) {
          victim = xchg(&ring_of(x, cq)[(v / kQCSpots) & (I::kQueueEntries - 1)], qc_tag(v / kQCSpots));
          reclaim = chunk_ok(x.m, victim);
        }
      }
      ring_put(x.m, x.base_ring, reclaim, victim);
    } else {
      // the directory, else the VL lockstep walk from back_ptr_
      uint32_t cur = want ? vl_dir_find(x, cq, v - v % kQCSpots, true) : kMarker;
      const bool direct = want && cur != kMarker;
      if (want && !direct) cur = ld(&cq->back_qc);
      bool todo = want && chunk_ok(x.m, cur);
      if (want && !todo) set_error(x.m, ERR_VSTART);
      bool old_want = false;
      uint32_t old_free = 0, mine = kMarker;
      Spin it(&x.m.error);
      while (ballot(todo)) {
        if (todo) {
          OURO_DIAG_ADD(9, 1);
          QueueChunkHdr* h = hdr<QueueChunkHdr>(x.m, cur);
          if (direct || ld(&h->virtual_start) / kQCSpots == v / kQCSpots) {
            if (v % kQCSpots == 0 && qc != kMarker) xchg(&h->next, qc);
            if constexpr (is_chunk_kind(K)) xchg(&qc_slots(x.m, cur)[v % kQCSpots], val);
            else st(&qc_slots(x.m, cur)[v % kQCSpots], val);  // a fresh queue chunk: the slot is MARKER
            mine = cur;
            todo = false;
          }
        }
        // one step per round for every lane: a lane whose next link is not set yet waits at the
        // round, not in a loop of its own, so the wave's other lanes (which may set the links that
        // another wave is waiting for, in another class queue) keep moving
        bool waiting = false;
        if (todo) {
          const uint32_t nx = ld(&hdr<QueueChunkHdr>(x.m, cur)->next);
          if (nx == kMarker) {
            waiting = true;
            if (it.expired()) {
              OURO_TIMEOUT(x.m);
              todo = false;
            }
          } else if (!chunk_ok(x.m, nx)) {
            set_error(x.m, ERR_VSTART);
            todo = false;
          } else {
            cur = nx;
          }
        }
        if (ballot(waiting) == ballot(todo)) backoff(it.it);  // nobody moved: back off
      }
      // every slot and link of the wave is written: count per queue chunk, then move back_ptr_ /
      // front_ptr_ over a queue chunk this completed
      if constexpr (!is_chunk_kind(K)) order();  // plain slot stores complete before the count
      const uint64_t grp = match<G>(mine != kMarker ? (uint64_t)mine : ~0ull);
      if (mine != kMarker && (int)lane() == first(grp)) {
        const uint32_t n = popc(grp);
        const uint32_t old = add(&hdr<QueueChunkHdr>(x.m, mine)->count, n * 0x10001u);
        if ((old & 0xFFFFu) + n == kQCSpots) vl_set_back(x, cq, mine);
        if (qc_empty_after_enqueue(old, n * 0x10001u)) {
          old_free = vl_set_front(x, cq, mine);
          old_want = true;
        }
      }
      vl_set_old(x, cq, old_want, old_free);
    }
  }
}

// take_lanes: every lane with want=true removes the element at virtual position v (spinning until
// it has been published). PageQueue::dequeue (PageQueue_impl.cuh:84-104), PageQueueVA::allocPage
// (:202-242), QueueChunk::dequeue (QueueChunk_impl.cuh:323-382). Queue-chunk counters as put_lanes.
template <int G, class I>
__device__ inline uint32_t take_lanes(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t v) {
  constexpr Kind K = I::kKind;
  uint32_t val = kMarker;
  if constexpr (is_std(K)) {
    if (want) val = ring_slot_take(x.m, &ring_of(x, cq)[v & (I::kQueueEntries - 1)]);
  } else {
    uint32_t c = kMarker;
    if (want) {
      if constexpr (is_va(K)) {
        c = va_wait_qc(x, cq, v);
      } else {
        c = vl_dir_find(x, cq, v - v % kQCSpots, true);
        if (c == kMarker) c = vl_locate(x, ld(&cq->front_qc), v);
      }
      if (c != kMarker) val = ring_slot_take(x.m, &qc_slots(x.m, c)[v % kQCSpots]);
    }
    // the slots are emptied before the queue-chunk count, whose last update may recycle it
    order();
    const uint64_t grp = match<G>(c != kMarker ? (uint64_t)c : ~0ull);
    const bool lead = c != kMarker && (int)lane() == first(grp);
    uint32_t old = 0, taken = 0;
    if (lead) {
      taken = popc(grp) << 16;
      old = sub(&hdr<QueueChunkHdr>(x.m, c)->count, taken);
    }
    const bool emptied = lead && old - taken == kQCSpots;  // QueueChunk::checkChunkEmptyDequeue
    if constexpr (is_va(K)) {
      uint32_t victim = 0;
      bool reclaim = false;
      if (emptied) {
        victim = xchg(&ring_of(x, cq)[(v / kQCSpots) & (I::kQueueEntries - 1)], qc_tag(v / kQCSpots));
        reclaim = chunk_ok(x.m, victim);
      }
      ring_put(x.m, x.base_ring, reclaim, victim);
    } else {
      const uint32_t old_free = emptied ? vl_set_front(x, cq, c) : 0u;
      vl_set_old(x, cq, emptied, old_free);
    }
  }
  return val;
}

// ------------------------------------------------------------------ semaphore
// BulkSemaphore::wait (BulkSemaphore_impl.cuh:199-300) for one group, restated on the debt line
// (sem_share): the group leader claims its pages with one fetch-add and learns from the displaced
// word how many chunks it must carve; it promises them (expected += k * ppc) and the carves then
// run wave-cooperatively, one leader at a time in lane order. No claim waits here: a page-queue
// claim waits on its own queue slot, a chunk-queue claim on its scan. The reference decides with a
// CAS loop on the word and parks waiters in `reserved`, which a cold class with thousands of
// concurrent groups turns into a serial chain of failed CASes.
template <class I, class Carve>
__device__ inline void sem_claim_n(Ctx<I>& x, ClassQueue* cq, bool lead, uint32_t num, Carve&& carve) {
  uint32_t k = 0;
  if (lead) {
    OURO_DIAG_ADD(12, 1);
    const uint64_t before = add(&cq->sem, (uint64_t)(-(int64_t)num));
    // pages per chunk only matter in debt; loading it unconditionally puts a plain read of the
    // semaphore's cache line next to every claim
    if (sem_debt(before) + (int64_t)num > 0) k = sem_share(before, (int64_t)num, ppc_of(x, cq));
    if (k) {
      OURO_DIAG_ADD(13, k);
      // a wrapped expected field would only make later claims over-carve, never under-carve
      add(&cq->sem, ((uint64_t)k * (uint64_t)ppc_of(x, cq)) << Sem::kExpShift);
    }
  }
  uint64_t carvers = ballot(k > 0);
  const uint32_t nact = popc(active());
  while (carvers) {
    const int l = first(carvers);
    carvers &= carvers - 1;
    ClassQueue* lq = reinterpret_cast<ClassQueue*>(readlane64(reinterpret_cast<uint64_t>(cq), l));
    const uint32_t kl = readlane(k, l);
    // a carve hands one chunk to each active lane: larger promises go in batches
    for (uint32_t d = 0; d < kl; d += nact) carve(lq, kl - d < nact ? kl - d : nact);
  }
}

template <class I, class Carve>
__device__ inline void sem_claim(Ctx<I>& x, ClassQueue* cq, uint64_t grp, bool want, Carve&& carve) {
  sem_claim_n(x, cq, want && (int)lane() == leader_of(grp), popc(grp), carve);
}

// Chunk-queue claims that cannot be served because the pool ran out give their pages back.
__device__ __forceinline__ bool pool_exhausted(const Memory& m) {
  return (__hip_atomic_load(&m.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & (ERR_OOM | ERR_QUEUE_FULL)) != 0;
}

// signalExpected (BulkSemaphore_impl.cuh:305-308) for n pages carved out of k_req * ppc expected.
__device__ __forceinline__ void sem_signal_expected(ClassQueue* cq, uint32_t n, uint32_t expected_removed) {
  add(&cq->sem, (uint64_t)n - ((uint64_t)expected_removed << Sem::kExpShift));
}
__device__ __forceinline__ void sem_signal(ClassQueue* cq, uint32_t n) { add(&cq->sem, (uint64_t)n); }
// Positions for a queue operation: lanes grouped by queue (match<G>), one fetch-add per group,
// rank order inside the group (atomicAggInc, Utility.cuh:76-95, when G = 64).
template <int G>
__device__ __forceinline__ uint32_t fetch_positions(uint32_t* counter_of_lane, bool want, uint32_t& n_out) {
  uint32_t pos = 0;
  const uint64_t key = want ? reinterpret_cast<uint64_t>(counter_of_lane) : ~0ull;
  const uint64_t grp = match<G>(key);
  const int L = first(grp);
  if (want && (int)lane() == L) pos = add(counter_of_lane, popc(grp));
  pos = shfl(pos, L) + rank_in(grp);
  n_out = popc(grp);
  return pos;
}

// ------------------------------------------------------------------ page-based instance
// Carve k chunks for class queue cq (PageQueue::allocPage's allocation lambda,
// PageQueue_impl.cuh:108-133, and enqueueChunk, :58-79), all active lanes cooperating.
template <class I>
__device__ inline uint32_t carve_pages(Ctx<I>& x, ClassQueue* cq, uint32_t k) {
  const uint64_t act = active();
  const uint32_t me = rank_in(act), nact = popc(act);
  const int l0 = first(act);
  const uint32_t ppc = ppc_of(x, cq), psize = page_size_of(x, cq);
  uint32_t chunk = kMarker;
  uint32_t k_ok = alloc_chunks(x.m, x.inst.reuse, I::kAdd, k, chunk);
  // capacity guard: never carve more pages into a class than its queue can hold
  uint32_t cap_ok = 1;
  if ((int)lane() == l0 && k_ok) {
    const uint64_t cap = is_vl(I::kKind) ? 0xFFFFFFFFull
                         : is_va(I::kKind) ? (uint64_t)I::kQueueEntries * kQCSpots - 2ull * kQCSpots
                                            : (uint64_t)I::kQueueEntries;
    const uint32_t before = add(&cq->carved, k_ok * ppc);
    if ((uint64_t)before + (uint64_t)k_ok * ppc > cap) {
      sub(&cq->carved, k_ok * ppc);
      set_error(x.m, ERR_QUEUE_FULL);
      cap_ok = 0;
    }
  }
  if (!readlane(cap_ok, l0)) k_ok = 0;
  if (me < k_ok) {
    PageChunkHdr* h = hdr<PageChunkHdr>(x.m, chunk);
    st(&h->page_size, psize);
    st(&h->number_pages, ppc);
    mark_chunk_class(x.m, chunk, I::kAdd, psize);
  }
  order();
  // every promised page gets a queue slot: the carved pages, then a null entry for each page of a
  // chunk the pool could not supply, so claims already counted on them are answered with nullptr
  const uint32_t n = k * ppc, n_real = k_ok * ppc;
  uint32_t v0 = 0;
  if ((int)lane() == l0 && n) v0 = add(&cq->back, n);
  v0 = readlane(v0, l0);
  for (uint32_t e0 = 0; e0 < n; e0 += nact) {
    const uint32_t e = e0 + me;
    const uint32_t j = (e < n_real ? e : (n_real ? n_real - 1 : 0)) / ppc;
    const uint32_t c = shfl(chunk, lane_of_rank(act, j));
    put_lanes<64>(x, cq, e < n, v0 + e, e < n_real ? make_index(c, e % ppc) : kNullPage);
  }
  if ((int)lane() == l0) sem_signal_expected(cq, n, k * ppc);
  return k_ok;
}

__device__ __forceinline__ uint32_t page_ppc_for(uint32_t chunk_size, uint32_t page_size) { return chunk_size / page_size; }

template <int W, class I>
__device__ inline void* alloc_pages(Ctx<I>& x, uint32_t qi, bool want) {
  constexpr int GS = W == 1 ? 64 : W;  // semaphore + front/back grouping
  ClassQueue* cq = &x.inst.q[want ? qi : 0];
  const uint64_t key = want ? (uint64_t)qi : ~0ull;
  const uint64_t grp = match<GS>(key);
  // the group's front_ positions (atomicAggInc) are fetched together with its semaphore claim: the
  // two atomics are in flight at once, and the positions do not depend on the claim's outcome (a
  // carve the claim owes fills the queue at back_, before any lane waits on its slot)
  const int L = first(grp);
  uint32_t pos = 0;
  if (want && (int)lane() == L) pos = add(&cq->front, popc(grp));
  {
    auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_pages(x, lq, k); };
    sem_claim(x, cq, grp, want, carve);
  }
  const uint32_t v = shfl(pos, L) + rank_in(grp);
  const uint32_t idx = take_lanes<W>(x, cq, want, v);
  if (!want || idx == kMarker || idx == kNullPage) return nullptr;
  if (!chunk_ok(x.m, index_chunk(idx)) || index_page(idx) >= ppc_of(x, cq)) {
    set_error(x.m, ERR_CORRUPT);
    return nullptr;
  }
  return chunk_addr(x.m, index_chunk(idx)) + kMeta + (uint64_t)index_page(idx) * page_size_of(x, cq);
}

template <int W, class I>
__device__ inline void free_pages(Ctx<I>& x, uint32_t qi, bool want, uint32_t chunk, uint32_t page) {
  constexpr int GS = W == 1 ? 64 : W;
  ClassQueue* cq = &x.inst.q[want ? qi : 0];
  uint32_t n;
  const uint32_t v = fetch_positions<GS>(&cq->back, want, n);
  put_lanes<W>(x, cq, want, v, make_index(chunk, page));
  // publish, then signal (the reference signals first, PageQueue_impl.cuh:24-31)
  if constexpr (W == 1) {
    if (want) sem_signal(cq, 1);
  } else {
    const uint64_t key = want ? (uint64_t)qi : ~0ull;
    const uint64_t grp = match<W>(key);
    if (want && (int)lane() == first(grp)) sem_signal(cq, popc(grp));
  }
}

// ------------------------------------------------------------------ chunk-based instance
__device__ __forceinline__ uint32_t mask_words(uint32_t ppc) {
  const uint32_t w = (ppc + 63u) / 64u;
  return w == 0 ? 1u : (w > kMaskWords ? kMaskWords : w);
}

// Enqueue chunk indices (ChunkQueue::enqueue, ChunkQueue_impl.cuh:20-38; ChunkQueueVA::enqueue,
// :344-376; ChunkQueueVL::enqueue, :264-273): count_ / back_ per queue group, queue_pos recorded.
// queue_pos names the position only once the put is complete (slot written, queue-chunk count
// updated): until then it reads enqueue-in-flight, so whoever removes the chunk by its queue_pos
// (a flash, a front relocation) never tombstones a slot before the put has written it.
template <int G, class I>
__device__ inline void enqueue_chunks(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t chunk, uint32_t gen) {
  uint32_t n;
  // count_ (VL keeps no count_ in the reference, but tracking it costs nothing here)
  const uint64_t key = want ? reinterpret_cast<uint64_t>(cq) : ~0ull;
  const uint64_t grp = match<G>(key);
  const int L = first(grp);
  if (want && (int)lane() == L) add(&cq->count, (int32_t)popc(grp));
  const uint32_t v = fetch_positions<G>(&cq->back, want, n);
  put_lanes<G>(x, cq, want, v, chunk);
  order();
  if (want) st(&hdr<ChunkHdr>(x.m, chunk)->queue_pos, qp_at(gen, v));
}

template <class I>
__device__ inline uint32_t carve_chunks(Ctx<I>& x, ClassQueue* cq, uint32_t k) {
  const uint64_t act = active();
  const uint32_t me = rank_in(act);
  const int l0 = first(act);
  const uint32_t ppc = ppc_of(x, cq), psize = page_size_of(x, cq);
  uint32_t chunk = kMarker;
  uint32_t k_ok = alloc_chunks(x.m, x.inst.reuse, I::kAdd, k, chunk);
  uint32_t cap_ok = 1;
  if ((int)lane() == l0 && k_ok) {
    const uint64_t cap = is_vl(I::kKind) ? 0xFFFFFFFFull
                         : is_va(I::kKind) ? (uint64_t)I::kQueueEntries * kQCSpots - 2ull * kQCSpots
                                            : (uint64_t)I::kQueueEntries;
    const uint32_t before = add(&cq->carved, k_ok);
    if ((uint64_t)before + k_ok > cap) {
      sub(&cq->carved, k_ok);
      set_error(x.m, ERR_QUEUE_FULL);
      cap_ok = 0;
    }
  }
  if (!readlane(cap_ok, l0)) k_ok = 0;
  // one generation per chunk carved (the carve counter advances by the chunks of this carve)
  uint32_t g0 = 0;
  if ((int)lane() == l0 && k_ok) g0 = add(&x.m.carves, k_ok);
  const uint32_t gen = make_gen(readlane(g0, l0) + me, (uint32_t)(cq - x.inst.q));
  if (me < k_ok) {
    // ChunkIndexChunk::initializeChunk with all pages free (ChunkAccess ctor, ChunkAccess.cuh:69-89)
    ChunkHdr* h = hdr<ChunkHdr>(x.m, chunk);
    mark_chunk_class(x.m, chunk, I::kAdd, psize);
    st(&h->page_size, psize);
    st(&h->size, ppc);
    st(count_gen(h), make_cg((int32_t)ppc, gen));
    st(&h->queue_pos, qp_pending(gen));  // enqueued below
    for (uint32_t w = 0; w < kMaskWords; ++w) {
      const uint32_t lo = w * 64u;
      uint64_t bits = 0;
      if (ppc > lo) bits = (ppc - lo >= 64u) ? ~0ull : ((1ull << (ppc - lo)) - 1ull);
      st(&h->mask[w], bits);
    }
  }
  order();
  // chunks enqueue in lane order: one group, positions = active ranks
  enqueue_chunks<64>(x, cq, me < k_ok, chunk, gen);
  if ((int)lane() == l0) sem_signal_expected(cq, k_ok * ppc, k * ppc);
  return k_ok;
}

// ---- chunk-queue scanning
// Peeks load kPeek consecutive slots (and then their chunk headers) at once: one round trip per
// batch instead of one per slot.

// Slots of the queue chunk holding virtual position v (VA / VL), from v on: nullptr when there is
// none to read, with `fill` what its slots read as (a reclaimed or bypassed queue chunk reads as
// tombstones, an unpublished one as markers). VL walks from `cursor` (front_ptr_ if unset) and
// leaves it at the queue chunk found. `dead` (optional) receives the number of positions from v to
// the end of its queue chunk when that queue chunk holds no live entry (reclaimed, bypassed, or
// every entry written so far already removed: its count's upper half, enqueued minus dequeued, is
// zero), else 0: a scan passes such a stretch in one step instead of slot by slot. (A put whose
// slot is still being written may land behind the scan; the scan finds it on its next pass.)
template <class I>
__device__ inline const uint32_t* slots_at(Ctx<I>& x, ClassQueue* cq, uint32_t v, uint32_t& cursor, uint32_t& fill,
                                          uint32_t* qc_out = nullptr, uint32_t* dead = nullptr) {
  constexpr Kind K = I::kKind;
  static_assert(!is_std(K), "standard queues index their ring directly");
  fill = kMarker;
  const uint32_t rest = kQCSpots - v % kQCSpots;
  if (dead) *dead = 0;
  if constexpr (is_va(K)) {
    const uint32_t e = ld(&ring_of(x, cq)[(v / kQCSpots) & (I::kQueueEntries - 1)]);
    if (e == qc_tag(v / kQCSpots)) {
      fill = kTomb;
      if (dead) *dead = rest;
    } else if (chunk_ok(x.m, e)) {
      const QueueChunkHdr* h = hdr<QueueChunkHdr>(x.m, e);
      const uint32_t vs = ld(&h->virtual_start), cnt = ld(&h->count);
      if (vs / kQCSpots == v / kQCSpots) {
        if (qc_out) *qc_out = e;
        if (dead && (cnt >> 16) == 0) *dead = rest;
        return qc_slots(x.m, e) + v % kQCSpots;
      }
    }
    return nullptr;
  } else {
    // the directory (its queue chunk may have been reclaimed and reused since: check the virtual
    // start the queue chunk itself holds), else the walk from the cursor
    const uint32_t vb = v - v % kQCSpots;
    if (const uint32_t d = vl_dir_find(x, cq, vb, false); d != kMarker) {
      const QueueChunkHdr* h = hdr<QueueChunkHdr>(x.m, d);
      const uint32_t vs = ld(&h->virtual_start), cnt = ld(&h->count);
      if (vs == vb) {
        cursor = d;
        if (qc_out) *qc_out = d;
        if (dead && (cnt >> 16) == 0) *dead = rest;
        return qc_slots(x.m, d) + v % kQCSpots;
      }
    }
    if (cursor == kMarker) cursor = ld(&cq->front_qc);
    uint32_t c = vl_locate(x, cursor, v);
    if (c == kMarker) {  // the cursor's queue chunk is already past v: retry from front_ptr_
      cursor = ld(&cq->front_qc);
      c = vl_locate(x, cursor, v);
    }
    if (c != kMarker) {
      cursor = c;
      if (qc_out) *qc_out = c;
      if (dead && (ld(&hdr<QueueChunkHdr>(x.m, c)->count) >> 16) == 0) *dead = rest;
      return qc_slots(x.m, c) + v % kQCSpots;
    }
    fill = kTomb;
    if (dead) *dead = rest;
    return nullptr;
  }
}

// Load the slots of virtual positions [v, v + n), n <= kPeek, which lie in one queue chunk for
// VA / VL; entries past n read markers. (ChunkQueue scan, ChunkQueue_impl.cuh:105-145;
// QueueChunk::access, QueueChunk.cuh:111-114, and the linked walk of accessLinked.)
// `dead`: as slots_at (always 0 for standard rings).
template <class I>
__device__ inline void peek_slots(Ctx<I>& x, ClassQueue* cq, uint32_t v, uint32_t n, uint32_t& cursor,
                                  uint32_t (&s)[kPeek], uint32_t& dead) {
  dead = 0;
  if constexpr (is_std(I::kKind)) {
#pragma unroll
    for (uint32_t i = 0; i < kPeek; ++i) s[i] = i < n ? ld(&ring_of(x, cq)[(v + i) & (I::kQueueEntries - 1)]) : kMarker;
  } else {
    uint32_t fill;
    const uint32_t* sl = slots_at(x, cq, v, cursor, fill, nullptr, &dead);
#pragma unroll
    for (uint32_t i = 0; i < kPeek; ++i) s[i] = (i < n && sl) ? ld(&sl[i]) : fill;
  }
}

// Offset of the first position in [v, v + n) that holds a chunk of this class queued exactly there
// (queue_pos == position, in the generation read with its count) with free pages; n when there is
// none, with `adv` the positions the scan may pass (n, or the rest of a queue chunk that holds no
// live entry). The queue_pos test also rejects a stale slot read out of a queue chunk that was
// recycled meanwhile. `cnt` and `gen` are the chunk's {count, gen} word as read.
template <class I>
__device__ inline uint32_t find_chunk(Ctx<I>& x, ClassQueue* cq, uint32_t v, uint32_t n, uint32_t& cursor,
                                      uint32_t& chunk, int32_t& cnt, uint32_t& gen, uint32_t& adv) {
  uint32_t s[kPeek], dead = 0;
  peek_slots(x, cq, v, n, cursor, s, dead);
  adv = n > dead ? n : dead;
  if (dead) return n;
  uint64_t cg[kPeek];
  uint64_t qp[kPeek];
  uint32_t ps[kPeek];
#pragma unroll
  for (uint32_t i = 0; i < kPeek; ++i) {
    ChunkHdr* h = hdr<ChunkHdr>(x.m, chunk_ok(x.m, s[i]) ? s[i] : 0u);
    cg[i] = ld(count_gen(h));
    qp[i] = ld(&h->queue_pos);
    ps[i] = ld(&h->page_size);
  }
  const uint32_t psize = page_size_of(x, cq), qi = (uint32_t)(cq - x.inst.q);
#pragma unroll
  for (uint32_t i = 0; i < kPeek; ++i) {
    const uint32_t g = cg_gen(cg[i]);
    if (i < n && chunk_ok(x.m, s[i]) && cg_count(cg[i]) > 0 && qp[i] == qp_at(g, v + i) && ps[i] == psize &&
        gen_class(g) == qi) {
      chunk = s[i];
      cnt = cg_count(cg[i]);
      gen = g;
      return i;
    }
  }
  return n;
}

// Tombstone the slot of a removed chunk at virtual position p (ChunkQueue_impl.cuh:434 and
// QueueChunk::deleteElement, QueueChunk_impl.cuh:349-359, write DeletionMarker instead). A queue
// chunk whose 2048 positions are now all dequeued is handed back through `reclaim` (VA: its index;
// VL: 0x80000000 | the number of front-pointer moves).
template <class I>
__device__ inline void tomb_slot(Ctx<I>& x, ClassQueue* cq, uint32_t p, uint32_t& cursor, uint32_t& reclaim) {
  constexpr Kind K = I::kKind;
  if constexpr (is_std(K)) {
    xchg(&ring_of(x, cq)[p & (I::kQueueEntries - 1)], kTomb);
  } else if constexpr (is_va(K)) {
    const uint32_t id = (p / kQCSpots) & (I::kQueueEntries - 1);
    const uint32_t qc = ld(&ring_of(x, cq)[id]);
    if (!chunk_ok(x.m, qc) || ld(&hdr<QueueChunkHdr>(x.m, qc)->virtual_start) / kQCSpots != p / kQCSpots) {
      set_error(x.m, ERR_CORRUPT);  // a queued chunk's queue chunk is always published
      return;
    }
    xchg(&qc_slots(x.m, qc)[p % kQCSpots], kTomb);
    const uint32_t old = sub(&hdr<QueueChunkHdr>(x.m, qc)->count, 1u << 16);
    if (qc_empty_after_dequeue(old)) {
      const uint32_t victim = xchg(&ring_of(x, cq)[id], qc_tag(p / kQCSpots));
      if (chunk_ok(x.m, victim)) reclaim = victim;
    }
  } else {
    uint32_t c = vl_dir_find(x, cq, p - p % kQCSpots, false);
    if (c == kMarker && cursor != kMarker) c = vl_locate(x, cursor, p);
    if (c == kMarker) c = vl_locate(x, ld(&cq->front_qc), p);
    if (c == kMarker) {
      set_error(x.m, ERR_CORRUPT);
      return;
    }
    xchg(&qc_slots(x.m, c)[p % kQCSpots], kTomb);
    const uint32_t old = sub(&hdr<QueueChunkHdr>(x.m, c)->count, 1u << 16);
    if (qc_empty_after_dequeue(old)) reclaim = vl_set_front(x, cq, c) | 0x80000000u;
  }
}

// Finish deferred queue-chunk reclamation (lane order across the wave).
template <class I>
__device__ inline void finish_reclaim(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t reclaim) {
  constexpr Kind K = I::kKind;
  if constexpr (is_va(K)) {
    ring_put(x.m, x.base_ring, want && reclaim != kMarker, reclaim);
  } else if constexpr (is_vl(K)) {
    const bool w = want && reclaim != kMarker;
    vl_set_old(x, cq, w, w ? (reclaim & 0x7FFFFFFFu) : 0u);
  }
}

// Lanes that moved a chunk's free count from <= 0 to > 0 make sure the chunk is in its class queue
// (FIRST_FREE, ChunkQueue_impl.cuh:166-170, and RE_ENQUEUE_CHUNK, :423-428). The queue_pos
// not-queued -> pending CAS admits exactly one enqueue however many lanes or removers race.
template <int G, class I>
__device__ inline void ensure_in_queue(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t chunk, uint32_t gen) {
  bool won = false;
  if (want) won = cas(&hdr<ChunkHdr>(x.m, chunk)->queue_pos, qp_not(gen), qp_pending(gen)) == qp_not(gen);
  if (ballot(won)) enqueue_chunks<G>(x, cq, won, chunk, gen);
}

// Remove an exhausted chunk from its class queue, wherever it sits (the reference removes with
// atomicMax(front_, v + 1), ChunkQueue_impl.cuh:432, which can strand live chunks in front of v).
// `pos` is the queue position the caller found the chunk at: the CAS queue_pos pos -> not-queued admits
// one remover and only for that entry. A caller that took the chunk's last page from an entry that
// has since been removed must not remove the chunk's next entry: a free may already have re-enqueued
// it at a position whose slot is still being written. Then the slot is tombstoned; the caller
// compacts front_ once for the wave afterwards (compact_wave where the lanes reconverge: one
// wave-wide pass instead of a contended serial one per removing lane). Returns true if removed;
// `again` reports a free that raced in (the caller re-enqueues: the free saw the chunk still queued and
// left it to us).
template <class I>
__device__ inline bool cq_remove(Ctx<I>& x, ClassQueue* cq, uint32_t chunk, uint32_t gen, uint32_t pos,
                                 uint32_t& cursor, uint32_t& reclaim, bool& again) {
  again = false;
  OURO_DIAG_ADD(7, 1);
  ChunkHdr* h = hdr<ChunkHdr>(x.m, chunk);
  const uint32_t p = pos;
  if (cas(&h->queue_pos, qp_at(gen, p), qp_not(gen)) != qp_at(gen, p)) return false;
  OURO_DIAG_ADD(8, 1);
  uint32_t rc = kMarker;
  tomb_slot(x, cq, p, cursor, rc);
  if (rc != kMarker) {
    if (reclaim != kMarker) finish_reclaim(x, cq, true, reclaim);
    reclaim = rc;
  }
  sub(&cq->count, 1);
  order();
  again = ld(&h->count) > 0;
  return true;
}

// Front relocation. Claims start at spread positions, so a live chunk near front_ can be passed
// over while the chunks behind it are consumed and re-queued at back_; front_ cannot move past it
// and the window [front_, back_) grows round after round (blocks held across rounds, many
// concurrent block claims), lengthening every scan and, left alone, running into the queue's
// capacity. When the window is more than twice the class's queued chunks plus kRelocSlack
// positions, compaction moves the live chunk that blocks front_ to back_: it removes the entry
// where it sits (queue_pos CAS, tombstone, count_, as cq_remove) and queues the chunk again at back_
// if it still has free pages (a chunk whose last page a claim took meanwhile stays out of the queue
// until a free gives it pages, as after any removal). The chunk's pages stay counted by the
// semaphore throughout. At most kMaxRelocate moves per call. The slack is small: with 4096 positions,
// a class of few chunks (512 pages each) let the chunk its claims had left partly used at front_
// block compaction until the window held thousands of dead positions, the page cursor ran past
// back_ and folded into them, and the per-thread claims scanned them every round (measured: the
// 64-lane malloc of a 1M-request round grew from 3 ms to 80 ms over 13 rounds).
constexpr uint32_t kRelocSlack = 64;
constexpr uint32_t kMaxRelocate = 16;

template <class I>
__device__ inline bool relocate_front(Ctx<I>& x, ClassQueue* cq) {
  constexpr Kind K = I::kKind;
  const int L = first(active());
  const bool lead = (int)lane() == L;
  uint32_t chunk = kMarker, f = 0, gen = 0;
  if (lead) {
    f = ld(&cq->front);
    const uint32_t b = ld(&cq->back), len = b - f;
    const int32_t live = ld(&cq->count);
    if (len < 0x80000000u && len > 2u * (uint32_t)(live > 0 ? live : 0) + kRelocSlack) {
      OURO_DIAG_ADD(9, 1);
      uint32_t e = kMarker;
      if constexpr (is_std(K)) {
        e = ld(&ring_of(x, cq)[f & (I::kQueueEntries - 1)]);
      } else {
        uint32_t cursor = kMarker, fill = kMarker;
        const uint32_t* sl = slots_at(x, cq, f, cursor, fill);
        e = sl ? ld(sl) : fill;
      }
      if (chunk_ok(x.m, e)) {
        ChunkHdr* h = hdr<ChunkHdr>(x.m, e);
        gen = cg_gen(ld(count_gen(h)));
        if (ld(&h->queue_pos) == qp_at(gen, f) && ld(&h->page_size) == page_size_of(x, cq) &&
            gen_class(gen) == (uint32_t)(cq - x.inst.q))
          chunk = e;
      }
      if (chunk != kMarker) OURO_DIAG_ADD(10, 1);
      else OURO_DIAG_LOG(((unsigned long long)f << 32) | e);
    }
  }
  chunk = readlane(chunk, L);
  f = readlane(f, L);
  gen = readlane(gen, L);
  if (chunk == kMarker) return false;
  uint32_t cursor = kMarker, reclaim = kMarker;
  bool again = false, removed = false;
  if (lead) removed = cq_remove(x, cq, chunk, gen, f, cursor, reclaim, again);
  reclaim = readlane(reclaim, L);
  finish_reclaim(x, cq, lead && reclaim != kMarker, reclaim);
  ensure_in_queue<1>(x, cq, lead && removed && again, chunk, gen);
  return readlane((uint32_t)removed, L) != 0;
}

// Wave-wide compaction of a chunk queue's front_ (compact_wave_pass), with front relocation; a pass
// that stopped at its step limit inside a long dead stretch is followed by another (the same budget
// of kMaxRelocate more passes).
template <uint32_t kPer = 16, class I>
__device__ inline void compact_wave(Ctx<I>& x, ClassQueue* cq) {
  bool more = compact_wave_pass<kPer>(x, cq);
  for (uint32_t r = 0; r < kMaxRelocate; ++r) {
    if (!more && !relocate_front(x, cq)) break;
    more = compact_wave_pass<kPer>(x, cq);
  }
}

// ChunkAccess::allocPage (ChunkAccess_impl.cuh:38-135), per lane, with 64-bit mask words.
enum class AccessMode { SUCCESS, DEQUEUE, RE_ENQUEUE, CONTINUE, ERROR };

// With recycling on, a chunk can be recycled and carved again between the scan that found it and
// this claim: the page is then reserved with a CAS on {count, gen}, which fails on a later life.
template <class I>
__device__ inline AccessMode access_alloc_lane(Ctx<I>& x, ChunkHdr* h, uint32_t gen, uint32_t& page) {
  int32_t cur = 0;
  AccessMode mode = AccessMode::SUCCESS;
  Spin it(&x.m.error);
  if (x.m.recycle_min != kRecycleOff) {
    uint64_t w = ld(count_gen(h));
    while (true) {
      if (cg_gen(w) != gen || cg_count(w) <= 0) return AccessMode::CONTINUE;
      const uint64_t seen = cas(count_gen(h), w, w - 1u);  // count > 0: no borrow into gen
      if (seen == w) break;
      w = seen;
      if (it.expired()) { OURO_TIMEOUT(x.m); return AccessMode::ERROR; }
    }
    cur = cg_count(w);
  } else {
    while ((cur = sub(&h->count, 1)) <= 0) {
      if ((cur = add(&h->count, 1)) < 0) return AccessMode::CONTINUE;
      if (cur == 0) mode = AccessMode::RE_ENQUEUE;
      if (it.expired()) { OURO_TIMEOUT(x.m); return AccessMode::ERROR; }
    }
  }
  if (cur == 1) mode = AccessMode::DEQUEUE;
  const uint32_t nw = mask_words(ld(&h->size));
  const uint32_t offset = (threadIdx.x + blockIdx.x) % 64u;
  uint32_t wi = threadIdx.x;
  while (true) {
    ++wi;
    uint64_t* word = &h->mask[wi % nw];
    uint64_t cm = ld(word);
    while (cm) {
      const uint64_t rot = offset ? ((cm >> offset) | (cm << (64u - offset))) : cm;
      const uint32_t b = ((uint32_t)__builtin_ctzll(rot) + offset) % 64u;
      const uint64_t old = and_(word, ~(1ull << b));
      if ((old >> b) & 1ull) {
        page = (wi % nw) * 64u + b;
        return mode;
      }
      cm = old & ~(1ull << b);
    }
    if (it.expired()) { OURO_TIMEOUT(x.m); return AccessMode::ERROR; }
  }
}

// Scan positions run over [front_, back_): a position at or past back_, or behind front_, restarts
// at front_.
template <class I>
__device__ __forceinline__ bool scan_wrap(ClassQueue* cq, uint32_t& v, uint32_t& cursor, uint32_t b) {
  const uint32_t f = ld(&cq->front);
  if (v - b < 0x80000000u || v - f >= 0x80000000u) {
    OURO_DIAG_ADD(6, 1);
    v = f;
    cursor = kMarker;
    return true;
  }
  return false;
}

// Batch length at v: at most kPeek, not past back_, and inside one queue chunk for VA / VL.
template <class I>
__device__ __forceinline__ uint32_t batch_at(uint32_t v, uint32_t b) {
  uint32_t n = b - v;
  if (n >= 0x80000000u) n = 0;
  if (n > kPeek) n = kPeek;
  if constexpr (!is_std(I::kKind)) {
    const uint32_t room = kQCSpots - v % kQCSpots;
    if (n > room) n = room;
  }
  return n;
}

// Where a claim of chunk pages starts scanning: the per-class page cursor (ClassQueue::scan) is
// advanced by the pages claimed, and its old value read as a queue position (cursor / ppc, one
// position per chunk of ppc pages) and folded into the live window [front_, back_). Concurrent claims
// thus begin at disjoint chunks, in queue order. The mapping assumes full chunks; where chunks hold
// fewer free pages (blocks held across rounds) claims spill over into the chunks the next claims
// start at, so a claim that passes kRedrawAfter positions without taking a page draws a new start
// for what it still needs (at most kMaxRedraws times): the extra draws advance the cursor at the rate
// at which pages are really found, instead of letting the spill cascade from claim to claim.
// A cursor that front_ has passed (chunks held fewer free pages than the cursor counts on, so
// consumption outran it) restarts at front_: the claim starts there and the cursor is set to front_
// plus its pages, so consumption resumes at the oldest live chunks and front_ can be compacted; a
// cursor folded back into the window at some arbitrary position would leave old live chunks behind
// and let the window grow round after round. A cursor that ran past back_ (claims ahead of the
// chunks being queued) folds back into the window and keeps its value: moving it back would send
// the claims after it over positions already consumed. The bounds were measured on 4M-request rounds
// with half the blocks held (VACQ): 64 positions and 4 redraws let the spill run on (95 / 140 M ops/s
// at W = 1 / 64), 24 and 12 cut it (231 / 610); 16 positions made the standard ring's late claims
// redraw over its tombstoned stretches (CQ W = 64 mallocs 3x slower on 1M mixed requests).
constexpr uint32_t kRedrawAfter = 24;
constexpr uint32_t kMaxRedraws = 12;

__device__ __forceinline__ uint32_t fold_start(ClassQueue* cq, uint32_t h) {
  const uint32_t f = ld(&cq->front), b = ld(&cq->back);
  const uint32_t len = b - f;
  return (len == 0 || len >= 0x80000000u) ? f : (h - f < len ? h : f + (h - f) % len);
}
// Page cursor value p0 (the old value of a fetch-add of n pages): its start position, or front_
// after restarting the cursor there (p0 is then front_'s cursor value).
template <class I>
__device__ __forceinline__ uint32_t start_of(Ctx<I>& x, ClassQueue* cq, uint64_t& p0, uint32_t n) {
  const uint32_t sh = __builtin_ctz(ppc_of(x, cq));
  const uint32_t h = (uint32_t)(p0 >> sh);
  const uint32_t f = ld(&cq->front), b = ld(&cq->back);
  const uint32_t len = b - f;
  if (len == 0 || len >= 0x80000000u) return f;
  const uint32_t d = h - f;
  if (d < len) return h;
  if (d < 0x80000000u) return f + d % len;  // past back_
  const uint64_t pf = (uint64_t)f << sh;
  cas(&cq->scan, p0 + n, pf + n);  // a claim that moved the cursor meanwhile keeps its value
  p0 = pf;
  return f;
}
template <class I>
__device__ __forceinline__ uint32_t draw_start(Ctx<I>& x, ClassQueue* cq, uint32_t n) {
  uint64_t p = add(&cq->scan, (uint64_t)n);
  return start_of(x, cq, p, n);
}

// One leader's claim of `num` pages from a chunk-based class queue, as a state machine that is
// stepped once per wave iteration: the leader finds chunks kPeek slots at a time, reserves pages on
// the chunk count with a CAS and claims them from the chunk's 64-bit mask words. Each step that
// claims bits reports them as (chunk | word << 29, bits, r0): the pages for ranks r0, r0 + 1, ...
// of the group, lowest bit first. Where a leader starts comes from a per-class page cursor: it adds
// its page count and starts at queue position cursor / ppc, so concurrent leaders begin at disjoint
// chunks, in queue order, and the consumed region stays behind the cursor (the reference starts
// every thread at front_, ChunkQueue_impl.cuh:101-105). A cursor outside [front_, back_) folds back
// into the window. (ChunkQueue::allocPage, ChunkQueue_impl.cuh:88-151, and ChunkAccess::allocPage,
// ChunkAccess_impl.cuh:38-135, for a whole group at once.)
template <class I>
struct ChunkClaim {
  uint32_t num = 0, v = 0, cursor = kMarker, need = 0, assigned = 0, t_rem = 0, w = 0, c = kMarker, wraps = 0;
  uint32_t reclaim = kMarker, miss = 0, redraws = 0;
  bool gave_up = false, removed = false;
  Spin it;

  __device__ ChunkClaim(Ctx<I>& x, ClassQueue* cq, bool leader, uint32_t n) : it(&x.m.error) {
    if (!leader) return;
    num = need = n;
    v = draw_start(x, cq, n);
  }
  __device__ bool finished() const { return assigned >= num || gave_up; }

  // One step of an unfinished leader. Returns bc != kMarker when it claimed pages.
  __device__ void step(Ctx<I>& x, ClassQueue* cq, uint32_t& bc, uint64_t& bits, uint32_t& r0) {
    bc = kMarker;
    bits = 0;
    r0 = 0;
    OURO_DIAG_ADD(0, 1);
    uint32_t n = 0;
    if (t_rem == 0) {
      OURO_DIAG_ADD(1, 1);
      if (miss >= kRedrawAfter && redraws < kMaxRedraws) {
        v = draw_start(x, cq, need);
        cursor = kMarker;
        miss = 0;
        ++redraws;
      }
      const uint32_t b = ld(&cq->back);
      wraps += scan_wrap<I>(cq, v, cursor, b);
      n = batch_at<I>(v, b);
      if ((wraps >= 2 || n == 0) && pool_exhausted(x.m)) {
        // the pool ran out and a whole window pass found nothing: unserved ranks get nullptr
        sem_signal(cq, need);
        gave_up = true;
        return;
      }
      uint32_t s = kMarker, adv = n, g = 0;
      int32_t cnt = 0;
      const uint32_t i = n ? find_chunk(x, cq, v, n, cursor, s, cnt, g, adv) : 0u;
      if (i < n) {
        ChunkHdr* h = hdr<ChunkHdr>(x.m, s);
        uint32_t t = 0;
        while (cnt > 0) {
          t = (uint32_t)cnt < need ? (uint32_t)cnt : need;
          const uint64_t seen = cas(count_gen(h), make_cg(cnt, g), make_cg(cnt - (int32_t)t, g));
          if (seen == make_cg(cnt, g)) break;
          OURO_DIAG_ADD(3, 1);
          cnt = cg_gen(seen) == g ? cg_count(seen) : 0;  // recycled meanwhile: not this chunk any more
        }
        if (cnt > 0) {
          miss = 0;
          c = s;
          t_rem = t;
          need -= t;
          const uint32_t size = ld(&h->size);
          w = ((size - (uint32_t)cnt) / 64u) % mask_words(size);
          bool again = false;
          if (cnt == (int32_t)t && cq_remove(x, cq, s, g, v + i, cursor, reclaim, again)) {
            removed = true;
            if (again) ensure_in_queue<1>(x, cq, true, s, g);
          }
          v += cnt == (int32_t)t ? i + 1u : i;
        } else {
          OURO_DIAG_ADD(4, 1);
          v += i + 1u;  // drained meanwhile; its last taker removes it
          miss += i + 1u;
        }
      } else {
        OURO_DIAG_ADD(2, adv);
        v += adv;
        miss += adv;
        if (n == 0) backoff(it.it);
      }
    }
    if (t_rem > 0) {
      ChunkHdr* h = hdr<ChunkHdr>(x.m, c);
      const uint32_t nw = mask_words(ld(&h->size));
      const uint64_t mw = ld(&h->mask[w]);
      const uint64_t take = lowest_bits(mw, t_rem);
      if (take == 0) {
        w = (w + 1) % nw;
      } else {
        const uint64_t old = and_(&h->mask[w], ~take);
        const uint64_t gotb = old & take;
        if (gotb != take) OURO_DIAG_ADD(5, 1);
        bits = gotb;
        bc = c | (w << 29);  // word index rides in the top bits (chunk < 2^23)
        r0 = assigned;
        assigned += popc(gotb);
        t_rem -= popc(gotb);
        if (gotb == take && t_rem > 0) w = (w + 1) % nw;
      }
    }
    if (it.expired()) {
      OURO_TIMEOUT(x.m);
      gave_up = true;
    }
  }
};

template <int W, class I>
__device__ inline void* alloc_chunk_pages(Ctx<I>& x, uint32_t qi, bool want) {
  constexpr int GS = W == 1 ? 64 : W;
  ClassQueue* cq = &x.inst.q[want ? qi : 0];
  const uint64_t key = want ? (uint64_t)qi : ~0ull;
  const uint64_t grp = match<GS>(key);
  {
    auto carve = [&](ClassQueue* lq, uint32_t k) { return carve_chunks(x, lq, k); };
    sem_claim(x, cq, grp, want, carve);
  }
  const bool ok = want;
  uint32_t chunk = kMarker, page = 0;
  uint32_t reclaim = kMarker;
  bool got = false, removed = false;
  if constexpr (W == 1) {
    // the reference's per-thread protocol: every lane walks the queue on its own and probes each
    // chunk's count with a decrement. Lanes start at the class's page cursor (one add per semaphore
    // group, rank order inside it) instead of all at front_, which a million lanes turn into one
    // long contended walk.
    const int L = first(grp);
    uint64_t p0 = 0;
    if (ok && (int)lane() == L) {
      p0 = add(&cq->scan, (uint64_t)popc(grp));
      (void)start_of(x, cq, p0, popc(grp));  // restarts a cursor that has left the window
    }
    p0 = shfl64(p0, L);
    if (ok) {
      uint32_t v = fold_start(cq, (uint32_t)((p0 + rank_in(grp)) >> __builtin_ctz(ppc_of(x, cq))));
      uint32_t cursor = kMarker, wraps = 0, miss = 0, redraws = 0;
      Spin it(&x.m.error);
      while (true) {
        OURO_DIAG_ADD(0, 1);
        if (miss >= kRedrawAfter && redraws < kMaxRedraws) {
          v = draw_start(x, cq, 1u);
          cursor = kMarker;
          miss = 0;
          ++redraws;
        }
        const uint32_t b = ld(&cq->back);
        wraps += scan_wrap<I>(cq, v, cursor, b);
        const uint32_t n = batch_at<I>(v, b);  // never walk towards a position at back_
        if ((wraps >= 2 || n == 0) && pool_exhausted(x.m)) {
          sem_signal(cq, 1);  // the pool ran out: give the claim back, return nullptr
          break;
        }
        // the first chunk queued in [v, v + n) with free pages (kPeek slots per round trip; the
        // chunks are tried in queue order, as the reference's one-by-one walk does)
        uint32_t c = kMarker, adv = n, g = 0;
        int32_t cnt = 0;
        const uint32_t i = n ? find_chunk(x, cq, v, n, cursor, c, cnt, g, adv) : 0u;
        if (i < n) {
          v += i;
          ChunkHdr* h = hdr<ChunkHdr>(x.m, c);
          const AccessMode md = access_alloc_lane(x, h, g, page);
          if (md == AccessMode::ERROR) break;
          if (md != AccessMode::CONTINUE) {
            chunk = c;
            got = true;
            bool again = false;
            if (md == AccessMode::RE_ENQUEUE) ensure_in_queue<1>(x, cq, true, c, g);
            if (md == AccessMode::DEQUEUE && cq_remove(x, cq, c, g, v, cursor, reclaim, again)) {
              removed = true;
              if (again) ensure_in_queue<1>(x, cq, true, c, g);
            }
            break;
          }
          adv = 1;  // drained meanwhile: go on after it
        }
        v += adv;
        miss += adv;
        if (it.expired()) { OURO_TIMEOUT(x.m); break; }
        backoff(it.it > 64 ? it.it : 0);
      }
    }
  } else {
    // group protocol: the group leader claims for the whole group (ChunkClaim); every step's pages
    // are broadcast and the lanes take them in rank order
    const int L = first(grp);
    const bool leader = ok && (int)lane() == L;
    ChunkClaim<I> cl(x, cq, leader, popc(grp));
    bool done = !ok;
    while (ballot(!done)) {
      uint32_t bc = kMarker, r0 = 0;
      uint64_t bits = 0;
      if (leader && !cl.finished()) cl.step(x, cq, bc, bits, r0);
      const uint32_t bcl = shfl(bc, L);
      const uint64_t bl = shfl64(bits, L);
      const uint32_t r0l = shfl(r0, L);
      const uint32_t gul = shfl((uint32_t)cl.gave_up, L);
      if (!done) {
        const uint32_t r = rank_in(grp);
        if (bcl != kMarker && r >= r0l && r < r0l + popc(bl)) {
          chunk = bcl & 0x1FFFFFFFu;
          page = ((bcl >> 29) * 64u) + nth_set_bit(bl, r - r0l);
          got = true;
          done = true;
        } else if (gul) {
          done = true;  // the leader gave up: this lane gets nullptr
        }
      }
    }
    reclaim = cl.reclaim;
    removed = cl.removed;
  }
  // the wave compacts each class queue it removed chunks from, one queue at a time
  for (uint64_t pend = ballot(removed); pend;) {
    ClassQueue* q = reinterpret_cast<ClassQueue*>(readlane64(reinterpret_cast<uint64_t>(cq), first(pend)));
    compact_wave(x, q);
    pend &= ~ballot(cq == q);
  }
  finish_reclaim(x, cq, true, reclaim);
  if (!got) return nullptr;
  return chunk_addr(x.m, chunk) + kMeta + (uint64_t)page * page_size_of(x, cq);
}

// Take n published pages out of a class semaphore, only if that many are unclaimed (the semaphore
// never owes claims more than it held before): BulkSemaphore::tryReduce (BulkSemaphore_impl.cuh:7-17),
// with a CAS loop instead of a fetch-add and undo, so concurrent claims never see the dip.
__device__ inline bool sem_try_take(ClassQueue* cq, uint32_t n) {
  uint64_t v = ld(&cq->sem);
  while (Sem::count(v) >= (int64_t)n) {
    const uint64_t seen = cas(&cq->sem, v, v - (uint64_t)n);
    if (seen == v) return true;
    v = seen;
  }
  return false;
}

// Chunk recycling (ChunkQueue::freePage's flash, ChunkQueue_impl.cuh:163-199). A lane whose free made
// every page of its chunk free again (`want`) tries it when recycling is on and the class queue holds
// more than recycle_min chunks, counting its own chunk (`unqueued` = 1 when the free found the chunk
// exhausted, hence out of the queue: it is recycled before it would be queued again, where the
// reference's FIRST_FREE enqueue wins and never recycles a chunk of one page, ChunkAccess_impl.cuh:
// 19-23): take `reduce` pages (those of the chunk already signalled) out of the
// class semaphore, only if that many are unclaimed; move the chunk's count from full to 0 with a CAS
// on {count, gen} (a claim that reserved a page meanwhile makes it fail, and the pages go back); take
// the chunk out of its queue (queue_pos, tombstone, count_; an enqueue in flight is waited for); clear
// its class; hand it to the instance's reuse ring, where any class of the instance (or a queue chunk)
// can take it. The wave's recycled chunks enter the ring together, and only while a quarter of the
// ring stays free for the queue chunks that share it; otherwise they are put back as they were. The
// lanes go one at a time in lane order (flashes are rare, and a single wave then acts exactly as the
// oracle's sequential model); the wave compacts each queue it removed chunks from afterwards. Returns true for a recycled chunk.
// (The reference flashes only above a fill level of 838,860 queued chunks, which its configurations
// never reach, and disables it in the VA and VL chunk queues, ChunkQueueVA_impl.cuh:229 and
// ChunkQueueVL_impl.cuh:160-192; here one switch covers all three kinds.)
template <class I>
__device__ inline bool flash_chunks(Ctx<I>& x, ClassQueue* cq, bool want, uint32_t chunk, uint32_t gen, uint32_t size,
                                    uint32_t reduce, uint32_t unqueued) {
  const uint32_t lim = x.m.recycle_min;
  if (lim == kRecycleOff) return false;
  uint64_t todo = ballot(want);
  if (!todo) return false;
  ChunkHdr* h = hdr<ChunkHdr>(x.m, want ? chunk : 0u);
  bool flashed = false, removed = false;
  uint32_t reclaim = kMarker;
  while (todo) {
    const int l = first(todo);
    todo &= todo - 1;
    if ((int)lane() != l) continue;
    if ((int64_t)ld(&cq->count) + unqueued <= (int64_t)lim || !sem_try_take(cq, reduce)) continue;
    if (cas(count_gen(h), make_cg((int32_t)size, gen), make_cg(0, gen)) != make_cg((int32_t)size, gen)) {
      sem_signal(cq, reduce);  // a claim reserved a page meanwhile: tryFlashChunk failed
      continue;
    }
    flashed = true;
    Spin it(&x.m.error);
    while (true) {
      const uint64_t qp = ld(&h->queue_pos);
      if (qp_queued(qp)) {
        if (cas(&h->queue_pos, qp, qp_recycled(gen)) == qp) {
          uint32_t cursor = kMarker, rc = kMarker;
          tomb_slot(x, cq, (uint32_t)qp, cursor, rc);
          if (rc != kMarker) {
            if (reclaim != kMarker) finish_reclaim(x, cq, true, reclaim);
            reclaim = rc;
          }
          sub(&cq->count, 1);
          removed = true;
          break;
        }
      } else if (qp == qp_not(gen)) {
        if (cas(&h->queue_pos, qp, qp_recycled(gen)) == qp) break;
      }
      // qp_pending: an enqueue is in flight; it publishes the position next
      if (it.expired()) {
        OURO_TIMEOUT(x.m);
        break;
      }
      backoff(it.it);
    }
    sub(&cq->carved, 1u);
    clear_chunk_class(x.m, chunk, I::kAdd);
    st(&h->page_size, 0u);
  }
  order();
  // a quarter of the ring stays free for queue chunks, which share it and cannot be put back
  const bool put = ring_put(x.m, x.inst.reuse, flashed, chunk, false, x.inst.reuse.size / 4);
  const bool back = flashed && !put;
  if (back) {
    // no room in the ring: the chunk returns to its class with every page free
    const uint32_t psize = page_size_of(x, cq);
    mark_chunk_class(x.m, chunk, I::kAdd, psize);
    st(&h->page_size, psize);
    add(&cq->carved, 1u);
    st(count_gen(h), make_cg((int32_t)size, gen));
    st(&h->queue_pos, qp_not(gen));  // ensure_in_queue takes it from here
    order();
  }
  ensure_in_queue<1>(x, cq, back, chunk, gen);
  if (back) sem_signal(cq, reduce);
  for (uint64_t pend = ballot(removed); pend;) {
    ClassQueue* q = reinterpret_cast<ClassQueue*>(readlane64(reinterpret_cast<uint64_t>(cq), first(pend)));
    compact_wave(x, q);
    pend &= ~ballot(cq == q);
  }
  finish_reclaim(x, cq, reclaim != kMarker, reclaim);
  return put;
}

template <int W, class I>
__device__ inline void free_chunk_pages(Ctx<I>& x, uint32_t qi, bool want, uint32_t chunk, uint32_t page) {
  ClassQueue* cq = &x.inst.q[want ? qi : 0];
  ChunkHdr* h = hdr<ChunkHdr>(x.m, want ? chunk : 0u);
  // ChunkAccess::freePage (ChunkAccess_impl.cuh:7-25): set the bit, then count. Groups (W = 16 / 64)
  // set their bits with one OR per mask word and segment, carrying all of them; W = 1 one per page.
  if constexpr (W == 1) {
    if (want) or_(&h->mask[page / 64u], 1ull << (page % 64u));
  } else {
    const uint64_t key = want ? ((uint64_t)chunk << 3) | (page / 64u) : ~0ull;
    for (uint64_t todo = ballot(want); todo;) {
      const int l = first(todo);
      const uint64_t same = ballot(key == readlane64(key, l)) & todo & segment_of<W>(l);
      todo &= ~same;
      uint64_t bits = 0;
      for (uint64_t r = same; r; r &= r - 1) bits |= 1ull << (readlane(page, first(r)) % 64u);
      if ((int)lane() == l) or_(&h->mask[page / 64u], bits);
    }
  }
  order();
  const uint64_t key = want ? (uint64_t)chunk : ~0ull;
  const uint64_t grp = match<W>(key);  // W = 1: every lane on its own
  const int L = first(grp);
  const uint32_t n = popc(grp);
  const bool lead = want && (int)lane() == L;
  const bool recycle = x.m.recycle_min != kRecycleOff;
  bool crossing = false, full = false;
  uint32_t gen = 0, size = 0;
  if (lead) {
    // the generation is read while the group's pages still pin the chunk's life
    if (recycle) {
      gen = ld(&h->gen);
      size = ld(&h->size);
    }
    const int32_t old = add(&h->count, (int32_t)n);
    crossing = old <= 0 && old + (int32_t)n > 0;
    if (!recycle && crossing) gen = ld(&h->gen);
    full = recycle && old + (int32_t)n == (int32_t)size;
  }
  const bool gone = flash_chunks(x, cq, full, chunk, gen, size, size - n, crossing ? 1u : 0u);
  // back into the queue before the pages are signalled
  ensure_in_queue<W>(x, cq, crossing && !gone, chunk, gen);
  if (lead && !gone) sem_signal(cq, n);
}

}  // namespace detail
}  // namespace ouro
