// The six instantiations under test and their two-instance MultiOuro* chains, under the reference's
// names (include/InstanceDefinitions.cuh:1-34), so a kernel written against the reference's types
// compiles unchanged.
#pragma once
#include "ouroboros.hpp"

namespace ouro {

// Queue organisation tags (queues/page/*.cuh, queues/chunk/*.cuh). They only select a Kind.
template <class>
struct PageQueue;
template <class>
struct PageQueueVA;
template <class>
struct PageQueueVL;
template <class>
struct ChunkQueue;
template <class>
struct ChunkQueueVA;
template <class>
struct ChunkQueueVL;

// Chunk<8192> (device/Chunk.cuh:15-26): every instance of a chain shares this base chunk.
template <size_t ChunkSize>
struct Chunk {
  static constexpr size_t size_ = ChunkSize;
  static constexpr size_t meta_data_size_ = kMeta;
  static constexpr size_t size() { return meta_data_size_ + size_; }
  static_assert(ChunkSize == kBaseChunk, "the pool layout is built on 8 KiB base chunks");
};

template <template <class> class Q>
struct KindOf;
template <>
struct KindOf<PageQueue> { static constexpr Kind value = Kind::PQ; };
template <>
struct KindOf<PageQueueVA> { static constexpr Kind value = Kind::VAPQ; };
template <>
struct KindOf<PageQueueVL> { static constexpr Kind value = Kind::VLPQ; };
template <>
struct KindOf<ChunkQueue> { static constexpr Kind value = Kind::CQ; };
template <>
struct KindOf<ChunkQueueVA> { static constexpr Kind value = Kind::VACQ; };
template <>
struct KindOf<ChunkQueueVL> { static constexpr Kind value = Kind::VLCQ; };

// OuroborosPages / OuroborosChunks (device/Ouroboros.cuh:53-245)
template <template <class> class Q, class ChunkBase, unsigned SMALLEST, unsigned NQ>
using OuroborosPages = Instance<KindOf<Q>::value, SMALLEST, NQ>;
template <template <class> class Q, class ChunkBase, unsigned SMALLEST, unsigned NQ>
using OuroborosChunks = Instance<KindOf<Q>::value, SMALLEST, NQ>;

constexpr unsigned NUM_QUEUES = 10;          // Parameters.h:9
constexpr unsigned SMALLEST_PAGE_SIZE = 16;  // Parameters.h:17
constexpr unsigned CHUNK_SIZE = SMALLEST_PAGE_SIZE << (NUM_QUEUES - 1);

using PQ = OuroborosPages<PageQueue, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;
using CQ = OuroborosChunks<ChunkQueue, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;
using VAPQ = OuroborosPages<PageQueueVA, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;
using VACQ = OuroborosChunks<ChunkQueueVA, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;
using VLPQ = OuroborosPages<PageQueueVL, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;
using VLCQ = OuroborosChunks<ChunkQueueVL, Chunk<CHUNK_SIZE>, SMALLEST_PAGE_SIZE, NUM_QUEUES>;

using OuroPQ = Ouroboros<PQ>;
using OuroCQ = Ouroboros<CQ>;
using OuroVAPQ = Ouroboros<VAPQ>;
using OuroVACQ = Ouroboros<VACQ>;
using OuroVLPQ = Ouroboros<VLPQ>;
using OuroVLCQ = Ouroboros<VLCQ>;

// Instance 2 of every chain: 16 KiB and 32 KiB pages on 32 KiB chunks (4 base chunks).
template <template <class> class Q>
using Second = Instance<KindOf<Q>::value, SMALLEST_PAGE_SIZE << NUM_QUEUES, 2>;

using MultiOuroPQ = Ouroboros<PQ, Second<PageQueue>>;
using MultiOuroCQ = Ouroboros<CQ, Second<ChunkQueue>>;
using MultiOuroVAPQ = Ouroboros<VAPQ, Second<PageQueueVA>>;
using MultiOuroVACQ = Ouroboros<VACQ, Second<ChunkQueueVA>>;
using MultiOuroVLPQ = Ouroboros<VLPQ, Second<PageQueueVL>>;
using MultiOuroVLCQ = Ouroboros<VLCQ, Second<ChunkQueueVL>>;

// Pool geometries the batched form's tuner chooses among (no reference counterpart): a single
// instance of CLASSES classes whose largest page is the CHUNK-byte instance chunk (a multiple of the
// 8 KiB base chunk, at most 512 pages per chunk), and chains whose class queue arrays hold QCAP
// entries instead of the reference's capacity.
template <template <class> class Q, unsigned CHUNK, unsigned CLASSES, unsigned QCAP = 0>
using OuroGeometry = Ouroboros<Instance<KindOf<Q>::value, (CHUNK >> (CLASSES - 1)), CLASSES, QCAP>>;
template <template <class> class Q, unsigned QCAP>
using MultiOuroCapacity = Ouroboros<Instance<KindOf<Q>::value, SMALLEST_PAGE_SIZE, NUM_QUEUES, QCAP>,
                                    Instance<KindOf<Q>::value, SMALLEST_PAGE_SIZE << NUM_QUEUES, 2, QCAP>>;

}  // namespace ouro
