/*
 * Sequential host simulation of the Ouroboros queue logic -- the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see ouro_sim.h). It restates the reference's algorithms with plain,
 * single-threaded state. Every function names the reference lines it follows. Where the MI355X
 * device code deliberately differs from the reference (64-bit page masks, group claims, batched
 * carves, publish-before-signal), the restatement follows the documented device rule and says so.
 */
#include "ouro_sim.h"

#include <stdlib.h>
#include <string.h>

#define MARKER 0xFFFFFFFFu
#define STRIDE 8320u          /* Chunk<8192>::size() (Chunk.cuh:23-26) */
#define META 128u             /* CHUNK_METADATA_SIZE (Parameters.h:19) */
#define BASE_CHUNK 8192u
#define PAGE_BITS 9u          /* NUM_BITS_FOR_PAGE (Parameters.h:31) */
#define STD_ENTRIES (1u << 23) /* page_queue_size (Parameters.h:11) */
#define VA_ENTRIES 16384u     /* virtual_queue_size (Parameters.h:12) */
#define REUSE_ENTRIES 65536u  /* chunk_queue_size (Parameters.h:10) */
#define QC_SPOTS 2048u        /* QueueChunk::num_spots_ (QueueChunk.cuh:27) */
#define OLD_HOLD 10u          /* LARGEST_OLD_COUNT_VALUE (Parameters.h:13) */
#define QC_RESERVE 32u        /* pool chunks only queue chunks may take */
#define STATE_BYTES 8192u
#define BLOCK 4096u
/* a chunk's queue position is a 32-bit value held in 64 bits; the sentinels are never positions */
#define NOT_QUEUED 0xFFFFFFFFFFFFFFFFull
#define QUEUE_PENDING 0xFFFFFFFFFFFFFFFEull
#define RECYCLED 0xFFFFFFFFFFFFFFFDull /* a chunk handed back to the reuse ring */
#define NULLPAGE 0xFFFFFFFCu  /* page-queue entry for a page the pool could not supply */
#define TOMB 0xFFFFFFFDu      /* chunk-queue slot whose chunk was removed */
#define PEEK 8u               /* slots looked at per scan step */
#define MAXL 256              /* requests of one call: a wave's 64, or a block-wide tile of up to 4 per
                                 thread of a 64-thread workgroup */
#define COMPACT_STEPS 256u    /* CAS steps per call of the per-lane compaction */
#define WAVE_COMPACT_STEPS 8192u /* front moves per call of the wave-wide compaction */
#define REDRAW_AFTER 24u      /* positions a claim passes without a page before it draws a new start */
#define MAX_REDRAWS 12u
#define RELOC_SLACK 64u       /* front relocation: window longer than 2 x queued chunks + this */
#define MAX_RELOCATE 16u

enum { ERR_OOM = 1, ERR_QUEUE_FULL = 2, ERR_VSTART = 4, ERR_TIMEOUT = 8, ERR_BAD_FREE = 16, ERR_TOO_LARGE = 32,
       ERR_REUSE_FULL = 64, ERR_CORRUPT = 256 };
enum { K_PQ = 0, K_CQ = 1, K_VAPQ = 2, K_VACQ = 3, K_VLPQ = 4, K_VLCQ = 5 };

static int is_chunk_kind(int k) { return k == K_CQ || k == K_VACQ || k == K_VLCQ; }
static int is_va(int k) { return k == K_VAPQ || k == K_VACQ; }
static int is_vl(int k) { return k == K_VLPQ || k == K_VLCQ; }
static int is_std(int k) { return k == K_PQ || k == K_CQ; }
/* VA pointer-ring entry of a reclaimed queue chunk: all positions of virtual queue chunk vqc dequeued */
static uint32_t qtag(uint32_t vqc) { return 0x80000000u | (vqc & 0x1FFFFFu); }
static int is_qtag(uint32_t e) { return (e >> 21) == 0x400u; }

typedef struct {
  int type; /* 0 unused, 1 page chunk, 2 chunk-queue chunk, 3 queue chunk */
  uint32_t page_size, size;
  int32_t count;
  uint64_t queue_pos;
  uint64_t mask[8];
  /* queue chunk view */
  uint32_t qc_count, vstart, next;
  uint32_t* slots;
} chunk_t;

typedef struct {
  int64_t cnt;
  int64_t exp, res;
  uint32_t front, back, carved, old_count, front_qc, back_qc, old_qc;
  uint64_t scan; /* chunk queues: page cursor */
  int32_t count;
  uint32_t page_size, ppc, ring_size;
  uint32_t** blocks; /* std ring, allocated in 4096-entry blocks */
  uint32_t* va;      /* VA queue-chunk ring */
} cls_t;

typedef struct {
  uint32_t* ring;
  uint32_t size, front, back;
  int32_t count;
} reuse_t;

typedef struct {
  int kind;
  uint32_t smallest, nq, largest, chunk_size, add;
  reuse_t reuse;
  cls_t q[10];
} inst_t;

struct ouro_sim {
  int variant;
  uint64_t max_chunks, alloc_size;
  uint32_t next_free, error, qc_reserve, qc_next;
  int block_order; /* 1: groups in class order (block-wide calls of one wavefront) */
  int agg;         /* aggregation width of the per-thread chunk claims: 64 (groups) or 1 (per lane) */
  uint32_t recycle_min; /* chunk recycling: MARKER = off, else the queued-chunk threshold */
  uint32_t* locator;
  chunk_t* chunks;
  inst_t inst[2];
  int ninst;
};

/* ------------------------------------------------------------------ closed forms */
/* QueueIndex::getQueueIndex (Helper.cuh:42-45): max(ceil(log2 size) - log2 smallest, 0). */
int sim_class_of(uint32_t size, uint32_t smallest) {
  int l = 0, s = 0;
  while ((1u << s) < smallest) ++s;
  while (l < 32 && (1ull << l) < size) ++l;
  return l > s ? l - s : 0;
}
static uint32_t make_index(uint32_t c, uint32_t p) { return (c << PAGE_BITS) + p; }

/* BulkSemaphore word (BulkSemaphore.cuh:5-84) as the device packs it: count in the low 40 bits with a
 * bias of 2^39, expected (pages of promised carves) in the high 24 bits, no reserved field. */
uint64_t sim_sem_make(int64_t count, uint64_t expected) {
  return (uint64_t)(count + (1ll << 39)) + (expected << 40);
}
int64_t sim_sem_count(uint64_t v) { return (int64_t)(v % (1ull << 40)) - (1ll << 39); }
uint64_t sim_sem_expected(uint64_t v) { return v >> 40; }

/* The carve decision of a claim (BulkSemaphore_impl.cuh:236-263 as the device restates it): a claim of
 * num pages moves the class debt, claims - (published + promised pages), from d to d + num and carves
 * one chunk per multiple of ppc that the move crosses above zero. */
static int64_t chunks_for_debt(int64_t d, int64_t ppc) { return d <= 0 ? 0 : (d + ppc - 1) / ppc; }
uint32_t sim_sem_share(int64_t count, int64_t expected, int64_t num, int64_t ppc) {
  const int64_t d = -(count + expected);
  return (uint32_t)(chunks_for_debt(d + num, ppc) - chunks_for_debt(d, ppc));
}

/* ChunkLocator::getChunkIndex (ChunkLocator.cuh:29-51). */
uint32_t sim_locate(const uint32_t* bits, uint32_t nwords, uint32_t idx) {
  int64_t w = idx >> 5;
  uint32_t mask = (uint32_t)((2ull << (idx & 31u)) - 1ull);
  if ((uint32_t)w >= nwords) return MARKER;
  for (;;) {
    uint32_t b = bits[w] & mask;
    if (b) {
      int hi = 31;
      while (!((b >> hi) & 1u)) --hi;
      return (uint32_t)(w * 32 + hi);
    }
    if (w == 0) return MARKER;
    --w;
    mask = 0xFFFFFFFFu;
  }
}

/* Pool arithmetic (MemoryInitialization.cuh:156-166, initNew :42-58). */
static uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }
static uint64_t entries_of(int kind) { return is_std(kind) ? STD_ENTRIES : (is_va(kind) ? VA_ENTRIES : 0); }
void sim_layout(int variant, uint64_t pool_bytes, uint64_t* max_chunks, uint64_t* allocation_size) {
  const uint64_t qb = align_up(entries_of(variant) * 4ull, BASE_CHUNK);
  const uint64_t rb = align_up((uint64_t)REUSE_ENTRIES * 4ull, BASE_CHUNK);
  const uint64_t mm = (rb + qb * 10) + (rb + qb * 2);
  const uint64_t total_required = align_up(STATE_BYTES + mm, STRIDE);
  const uint64_t aligned = align_up(pool_bytes, STRIDE);
  const uint64_t mc = aligned > total_required ? (aligned - total_required) / STRIDE : 0;
  const uint64_t adjacency = align_up(mc * STRIDE, 128);
  const uint64_t loc = align_up((mc + 31) / 32 * 4, STRIDE);
  if (max_chunks) *max_chunks = mc;
  if (allocation_size) *allocation_size = align_up(total_required + adjacency + loc, STRIDE);
}

/* ------------------------------------------------------------------ storage helpers */
static uint32_t* std_slot(cls_t* q, uint32_t pos) {
  const uint32_t p = pos & (q->ring_size - 1);
  uint32_t** b = &q->blocks[p / BLOCK];
  if (!*b) {
    *b = (uint32_t*)malloc(BLOCK * 4);
    memset(*b, 0xFF, BLOCK * 4);
  }
  return &(*b)[p % BLOCK];
}
static chunk_t* ch(ouro_sim* s, uint32_t c) { return &s->chunks[c]; }
static int chunk_ok(const ouro_sim* s, uint32_t c) { return c < s->max_chunks; }

/* IndexQueue::enqueue (Queue_impl.cuh:20-38) */
static void ring_put(ouro_sim* s, reuse_t* r, uint32_t value) {
  const int32_t f = r->count++;
  if (f >= (int32_t)r->size) {
    r->count--;
    s->error |= ERR_REUSE_FULL;
    return;
  }
  const uint32_t pos = r->back++ & (r->size - 1);
  if (r->ring[pos] != MARKER) s->error |= ERR_TIMEOUT;
  r->ring[pos] = value;
}
/* IndexQueue::dequeue (Queue_impl.cuh:54-67); NO value -> MARKER */
static uint32_t ring_take(reuse_t* r) {
  if (r->count <= 0) return MARKER;
  r->count--;
  const uint32_t pos = r->front++ & (r->size - 1);
  const uint32_t v = r->ring[pos];
  r->ring[pos] = MARKER;
  return v;
}

/* allocateChunk<QUEUECHUNK> (Ouroboros.cuh:108-136) for k chunks: reuse ring first, then bump.
 * Validity follows the reference's predicate chunk + add < maxChunks (:133) below the queue-chunk
 * reserve (the top qc_reserve chunks of a virtualized pool); queue chunks that find the bump region
 * spent continue in the reserve. Returns the valid count. */
static uint32_t alloc_chunks(ouro_sim* s, reuse_t* r, uint32_t add, uint32_t k, uint32_t* out, int queue) {
  /* data carves leave the last qc_reserve / 2 recycled chunks of a virtualized pool to queue chunks */
  const int32_t keep = queue ? 0 : (int32_t)(s->qc_reserve / 2);
  uint32_t t = 0;
  while (t < k && r->count > keep) out[t++] = ring_take(r);
  uint32_t ok = 0;
  for (uint32_t j = 0; j < t; ++j)
    if (chunk_ok(s, out[j])) ++ok;
  const uint64_t limit = s->max_chunks - s->qc_reserve;
  if (k > t) {
    const uint32_t base = s->next_free;
    s->next_free += (k - t) * add;
    for (uint32_t j = t; j < k; ++j) {
      uint32_t c = base + (j - t) * add;
      int valid = (uint64_t)c + add < limit;
      if (!valid && queue) {
        const uint32_t idx = s->qc_next++;
        c = (uint32_t)limit + idx;
        valid = idx < s->qc_reserve && c < s->max_chunks;
      }
      out[j] = c;
      if (valid) {
        s->locator[c >> 5] |= 1u << (c & 31u);
        ++ok;
      }
    }
  }
  if (ok < k) s->error |= ERR_OOM;
  return ok;
}

/* QueueChunk::initializeChunk (QueueChunk.cuh:221-225) with cleaned slots. */
static void qc_init(ouro_sim* s, uint32_t c, uint32_t vstart) {
  chunk_t* q = ch(s, c);
  q->type = 3;
  q->page_size = 0;
  q->qc_count = 0;
  q->vstart = vstart;
  q->next = MARKER;
  if (!q->slots) q->slots = (uint32_t*)malloc(QC_SPOTS * 4);
  memset(q->slots, 0xFF, QC_SPOTS * 4);
}

/* ------------------------------------------------------------------ VL pointer maintenance */
/* QueueChunk::setBackPointer (QueueChunk_impl.cuh:440-467) */
static void vl_set_back(ouro_sim* s, cls_t* q, uint32_t c) {
  for (;;) {
    const uint32_t nx = ch(s, c)->next;
    if (!chunk_ok(s, nx) || q->back_qc != c) break;
    q->back_qc = nx;
    if ((ch(s, nx)->qc_count & 0xFFFFu) == QC_SPOTS) c = nx;
    else break;
  }
}
/* QueueChunk::setFrontPointer (QueueChunk_impl.cuh:471-495) */
static uint32_t vl_set_front(ouro_sim* s, cls_t* q, uint32_t c) {
  uint32_t moved = 0;
  for (;;) {
    const uint32_t nx = ch(s, c)->next;
    if (!chunk_ok(s, nx) || q->front_qc != c) break;
    q->front_qc = nx;
    ++moved;
    if (ch(s, nx)->qc_count == QC_SPOTS) c = nx;
    else break;
  }
  return moved;
}
/* QueueChunk::setOldPointer (QueueChunk_impl.cuh:499-551): hold back OLD_HOLD emptied queue chunks,
 * recycle the rest through the base reuse ring. */
static void vl_set_old(ouro_sim* s, cls_t* q, uint32_t free_count) {
  if (q->old_count < OLD_HOLD) {
    const uint32_t cur = q->old_count;
    q->old_count += free_count;
    if (cur < OLD_HOLD) {
      if (cur + free_count > OLD_HOLD) {
        free_count = cur + free_count - OLD_HOLD;
        q->old_count -= free_count;
      } else {
        free_count = 0;
      }
    }
  }
  while (free_count > 0) {
    const uint32_t cur = q->old_qc;
    const uint32_t nx = ch(s, cur)->next;
    if (!chunk_ok(s, nx)) {
      s->error |= ERR_TIMEOUT;
      break;
    }
    q->old_qc = nx;
    --free_count;
    ring_put(s, &s->inst[0].reuse, cur);
  }
}

/* QueueChunk::locateQueueChunkForPosition (QueueChunk_impl.cuh:386-414); MARKER when the walk
 * starts past v's queue chunk (that one is gone). */
static uint32_t vl_locate(ouro_sim* s, uint32_t c, uint32_t v) {
  const uint32_t vb = v - v % QC_SPOTS;
  for (;;) {
    if (!chunk_ok(s, c)) {
      s->error |= ERR_VSTART;
      return MARKER;
    }
    /* full 32-bit virtual starts: differences wrap with the positions */
    const uint32_t vs = ch(s, c)->vstart;
    if (vs == vb) return c;
    if (vs - vb < 0x80000000u) return MARKER;
    c = ch(s, c)->next;
  }
}

static void compact(ouro_sim* s, inst_t* I, cls_t* q, int wave);

/* ------------------------------------------------------------------ slot sequences */
/* Queue-chunk counters (QueueChunk count_) change once per queue chunk of a call, at the lowest lane
 * holding it (the device groups lanes by queue chunk with one atomic per group). group_size returns
 * the number of lanes holding qc[i] if i is the lowest of them, else 0. */
static int group_size(const uint32_t* qc, int n, int i) {
  if (qc[i] == MARKER) return 0;
  for (int j = 0; j < i; ++j)
    if (qc[j] == qc[i]) return 0;
  int m = 0;
  for (int j = i; j < n; ++j) m += qc[j] == qc[i];
  return m;
}

/* put for a set of lanes already sorted in lane order: positions v[i], values val[i].
 * Queue chunks needed by positions at a chunk boundary are allocated first, in lane order
 * (PageQueueVA_impl.cuh:196-229, QueueChunk_impl.cuh:137-185 with the device's pre-phase). */
static void put_lanes(ouro_sim* s, inst_t* I, cls_t** qs, int n, const uint32_t* v, const uint32_t* val) {
  const int K = I->kind;
  if (is_std(K)) {
    for (int i = 0; i < n; ++i) {
      uint32_t* slot = std_slot(qs[i], v[i]);
      if (*slot == TOMB && is_chunk_kind(K)) compact(s, I, qs[i], 0);
      if (*slot != MARKER) s->error |= ERR_TIMEOUT;
      else *slot = val[i];
    }
    return;
  }
  uint32_t qc[64];
  int need[64], nneed = 0;
  for (int i = 0; i < n; ++i) {
    qc[i] = MARKER;
    if (v[i] % QC_SPOTS == 0) need[nneed++] = i;
  }
  if (nneed) {
    uint32_t got[64];
    const uint32_t ok = alloc_chunks(s, &s->inst[0].reuse, 1, (uint32_t)nneed, got, 1);
    for (int j = 0; j < nneed; ++j)
      if ((uint32_t)j < ok) {
        qc[need[j]] = got[j];
        qc_init(s, got[j], v[need[j]] + QC_SPOTS);
        if (is_va(K)) {
          /* publish over a marker or a reclaimed tag only */
          uint32_t* e = &qs[need[j]]->va[((v[need[j]] / QC_SPOTS) + 1) & (qs[need[j]]->ring_size - 1)];
          if (*e == MARKER || is_qtag(*e)) *e = got[j];
          else s->error |= ERR_QUEUE_FULL;
        }
      }
  }
  uint32_t held[64];
  if (is_va(K)) {
    uint32_t victims[64];
    int nv = 0;
    for (int i = 0; i < n; ++i) {
      cls_t* q = qs[i];
      held[i] = MARKER;
      const uint32_t id = (v[i] / QC_SPOTS) & (q->ring_size - 1);
      const uint32_t c = q->va[id];
      if (!chunk_ok(s, c) || ch(s, c)->vstart / QC_SPOTS != v[i] / QC_SPOTS) {
        s->error |= ERR_VSTART;
        continue;
      }
      chunk_t* x = ch(s, c);
      if (x->slots[v[i] % QC_SPOTS] != MARKER) s->error |= ERR_TIMEOUT;
      x->slots[v[i] % QC_SPOTS] = val[i];
      held[i] = c;
    }
    for (int i = 0; i < n; ++i) {
      const int m = group_size(held, n, i);
      if (!m) continue;
      chunk_t* x = ch(s, held[i]);
      const uint32_t added = (uint32_t)m * 0x10001u, old = x->qc_count;
      x->qc_count += added;
      if (old + added == QC_SPOTS) {
        const uint32_t id = (v[i] / QC_SPOTS) & (qs[i]->ring_size - 1);
        victims[nv++] = qs[i]->va[id];
        qs[i]->va[id] = qtag(v[i] / QC_SPOTS);
      }
    }
    for (int i = 0; i < nv; ++i) ring_put(s, &s->inst[0].reuse, victims[i]);
    return;
  }
  /* VL: walk, link at local position 0, write; then the counters and the back / front pointers */
  uint32_t oldf[64];
  int oldw[64];
  for (int i = 0; i < n; ++i) {
    cls_t* q = qs[i];
    oldw[i] = 0;
    oldf[i] = 0;
    held[i] = vl_locate(s, q->back_qc, v[i]);
    if (held[i] == MARKER) continue;
    chunk_t* x = ch(s, held[i]);
    if (v[i] % QC_SPOTS == 0 && qc[i] != MARKER) x->next = qc[i];
    x->slots[v[i] % QC_SPOTS] = val[i];
  }
  for (int i = 0; i < n; ++i) {
    const int m = group_size(held, n, i);
    if (!m) continue;
    chunk_t* x = ch(s, held[i]);
    const uint32_t added = (uint32_t)m * 0x10001u, old = x->qc_count;
    x->qc_count += added;
    if ((old & 0xFFFFu) + (uint32_t)m == QC_SPOTS) vl_set_back(s, qs[i], held[i]);
    if (old + added == QC_SPOTS) {
      oldf[i] = vl_set_front(s, qs[i], held[i]);
      oldw[i] = 1;
    }
  }
  for (int i = 0; i < n; ++i)
    if (oldw[i]) vl_set_old(s, qs[i], oldf[i]);
}

/* take for lanes in lane order (PageQueue::dequeue PageQueue_impl.cuh:84-104, PageQueueVA::allocPage
 * :202-242, QueueChunk::dequeue QueueChunk_impl.cuh:323-382). */
static void take_lanes(ouro_sim* s, inst_t* I, cls_t** qs, int n, const uint32_t* v, uint32_t* out) {
  const int K = I->kind;
  if (is_std(K)) {
    for (int i = 0; i < n; ++i) {
      uint32_t* slot = std_slot(qs[i], v[i]);
      out[i] = *slot;
      if (*slot == MARKER) s->error |= ERR_TIMEOUT;
      *slot = MARKER;
    }
    return;
  }
  uint32_t held[64];
  for (int i = 0; i < n; ++i) {
    cls_t* q = qs[i];
    out[i] = MARKER;
    held[i] = MARKER;
    uint32_t c;
    if (is_va(K)) {
      c = q->va[(v[i] / QC_SPOTS) & (q->ring_size - 1)];
      if (!chunk_ok(s, c) || ch(s, c)->vstart / QC_SPOTS != v[i] / QC_SPOTS) {
        s->error |= ERR_VSTART;
        continue;
      }
    } else {
      c = vl_locate(s, q->front_qc, v[i]);
      if (c == MARKER) continue;
    }
    chunk_t* x = ch(s, c);
    out[i] = x->slots[v[i] % QC_SPOTS];
    if (out[i] == MARKER) s->error |= ERR_TIMEOUT;
    x->slots[v[i] % QC_SPOTS] = MARKER;
    held[i] = c;
  }
  /* counters per queue chunk; an emptied queue chunk is reclaimed (VA) or passed by front_ptr_ (VL) */
  uint32_t victims[64], oldf[64];
  int nv = 0, oldw[64];
  for (int i = 0; i < n; ++i) {
    oldw[i] = 0;
    oldf[i] = 0;
    const int m = group_size(held, n, i);
    if (!m) continue;
    chunk_t* x = ch(s, held[i]);
    const uint32_t taken = (uint32_t)m << 16, old = x->qc_count;
    x->qc_count -= taken;
    if (old - taken != QC_SPOTS) continue;
    if (is_va(K)) {
      const uint32_t id = (v[i] / QC_SPOTS) & (qs[i]->ring_size - 1);
      victims[nv++] = qs[i]->va[id];
      qs[i]->va[id] = qtag(v[i] / QC_SPOTS);
    } else {
      oldf[i] = vl_set_front(s, qs[i], held[i]);
      oldw[i] = 1;
    }
  }
  for (int i = 0; i < nv; ++i) ring_put(s, &s->inst[0].reuse, victims[i]);
  for (int i = 0; i < n; ++i)
    if (oldw[i]) vl_set_old(s, qs[i], oldf[i]);
}

/* ------------------------------------------------------------------ chunk-queue scanning */
/* Slots of positions [v, v + n), n <= PEEK, inside one queue chunk for VA / VL: a reclaimed or
 * bypassed queue chunk reads as tombstones, an unpublished one as markers (the device's peek_slots;
 * ChunkQueue_impl.cuh:105-145, QueueChunk.cuh:111-114). *dead (if given): the positions from v to
 * the end of its queue chunk when that queue chunk holds no live entry (reclaimed, bypassed, or its
 * count's enqueued-minus-dequeued half zero), else 0; a scan passes such a stretch in one step. */
static void peek_slots(ouro_sim* s, inst_t* I, cls_t* q, uint32_t v, uint32_t n, uint32_t* cursor,
                       uint32_t* out, uint32_t* dead) {
  const uint32_t rest = QC_SPOTS - v % QC_SPOTS;
  if (dead) *dead = 0;
  if (is_std(I->kind)) {
    for (uint32_t i = 0; i < PEEK; ++i) out[i] = i < n ? *std_slot(q, v + i) : MARKER;
    return;
  }
  const uint32_t* sl = NULL;
  uint32_t fill = MARKER;
  if (is_va(I->kind)) {
    const uint32_t e = q->va[(v / QC_SPOTS) & (q->ring_size - 1)];
    if (e == qtag(v / QC_SPOTS)) {
      fill = TOMB;
      if (dead) *dead = rest;
    } else if (chunk_ok(s, e) && ch(s, e)->slots && ch(s, e)->vstart / QC_SPOTS == v / QC_SPOTS) {
      sl = ch(s, e)->slots + v % QC_SPOTS;
      if (dead && (ch(s, e)->qc_count >> 16) == 0) *dead = rest;
    }
  } else {
    if (*cursor == MARKER) *cursor = q->front_qc;
    uint32_t c = vl_locate(s, *cursor, v);
    if (c == MARKER) { /* the cursor's queue chunk is already past v: retry from front_ptr_ */
      *cursor = q->front_qc;
      c = vl_locate(s, *cursor, v);
    }
    if (c != MARKER) {
      *cursor = c;
      sl = ch(s, c)->slots + v % QC_SPOTS;
      if (dead && (ch(s, c)->qc_count >> 16) == 0) *dead = rest;
    } else {
      fill = TOMB;
      if (dead) *dead = rest;
    }
  }
  for (uint32_t i = 0; i < PEEK; ++i) out[i] = (i < n && sl) ? sl[i] : fill;
}

/* First offset in [v, v + n) that holds a chunk of this class queued exactly there with free pages;
 * n when there is none, with *adv the positions the scan passes (n, or the rest of a queue chunk
 * that holds no live entry). */
static uint32_t find_chunk(ouro_sim* s, inst_t* I, cls_t* q, uint32_t v, uint32_t n, uint32_t* cursor,
                           uint32_t* chunk, uint32_t* adv) {
  uint32_t sl[PEEK], dead = 0;
  peek_slots(s, I, q, v, n, cursor, sl, &dead);
  *adv = n > dead ? n : dead;
  if (dead) return n;
  for (uint32_t i = 0; i < n; ++i) {
    if (!chunk_ok(s, sl[i])) continue;
    const chunk_t* x = ch(s, sl[i]);
    if (x->count > 0 && x->queue_pos == (uint64_t)(v + i) && x->page_size == q->page_size) {
      *chunk = sl[i];
      return i;
    }
  }
  return n;
}

/* Advance front over dead positions: tombstones, a VA queue chunk whose pointer-ring entry holds its
 * reclaimed tag, VL queue chunks front_ptr_ has already left. The per-lane form (wave = 0) takes at
 * most COMPACT_STEPS steps, a step on a standard ring passing up to PEEK leading tombstones at once;
 * the wave-wide form makes at most WAVE_COMPACT_STEPS moves, a move being one position or one jump
 * over a whole queue chunk. */
static int compact_pass(ouro_sim* s, inst_t* I, cls_t* q, int wave) {
  uint32_t f = q->front;
  const uint32_t b = q->back;
  if (is_std(I->kind)) {
    const uint32_t steps = wave ? WAVE_COMPACT_STEPS : COMPACT_STEPS, per = wave ? 1u : PEEK;
    uint32_t i = 0;
    for (; i < steps && f != b && b - f < 0x80000000u; ++i) {
      const uint32_t n = b - f < per ? b - f : per;
      uint32_t k = 0;
      while (k < n && *std_slot(q, f + k) == TOMB) ++k;
      if (k == 0) return 0;
      for (uint32_t j = 0; j < k; ++j) *std_slot(q, f + j) = MARKER;
      f += k;
      q->front = f;
      if (k < n) return 0;
    }
    return wave && i == steps; /* stopped at the step limit */
  }
  const uint32_t steps = wave ? WAVE_COMPACT_STEPS : COMPACT_STEPS;
  uint32_t i = 0;
  for (; i < steps && f != b && b - f < 0x80000000u; ++i) {
    uint32_t nf = f + 1u;
    if (is_va(I->kind)) {
      const uint32_t e = q->va[(f / QC_SPOTS) & (q->ring_size - 1)];
      if (e == qtag(f / QC_SPOTS)) {
        nf = (f / QC_SPOTS + 1u) * QC_SPOTS;
      } else if (!chunk_ok(s, e) || !ch(s, e)->slots || ch(s, e)->vstart / QC_SPOTS != f / QC_SPOTS ||
                 ch(s, e)->slots[f % QC_SPOTS] != TOMB) {
        return 0;
      }
    } else {
      const uint32_t c = q->front_qc;
      if (!chunk_ok(s, c)) return 0;
      const uint32_t vs = ch(s, c)->vstart, fb = f - f % QC_SPOTS;
      if (vs != fb) {
        if (vs - fb >= 0x80000000u) return 0;
        nf = vs;
      } else if (ch(s, c)->slots[f % QC_SPOTS] != TOMB) {
        return 0;
      }
    }
    q->front = nf;
    f = nf;
  }
  return wave && i == steps; /* stopped at the step limit */
}

/* Tombstone the slot of a removed chunk at p (ChunkQueue_impl.cuh:434, QueueChunk::deleteElement
 * QueueChunk_impl.cuh:349-359 write DeletionMarker). A queue chunk emptied here is reclaimed later. */
static void tomb_slot(ouro_sim* s, inst_t* I, cls_t* q, uint32_t p, uint32_t* cursor, uint32_t* reclaim) {
  if (is_std(I->kind)) {
    *std_slot(q, p) = TOMB;
    return;
  }
  if (is_va(I->kind)) {
    const uint32_t id = (p / QC_SPOTS) & (q->ring_size - 1);
    const uint32_t c = q->va[id];
    if (!chunk_ok(s, c) || ch(s, c)->vstart / QC_SPOTS != p / QC_SPOTS) {
      s->error |= ERR_CORRUPT;
      return;
    }
    chunk_t* x = ch(s, c);
    x->slots[p % QC_SPOTS] = TOMB;
    const uint32_t old = x->qc_count;
    x->qc_count -= 0x10000u;
    if (old - 0x10000u == QC_SPOTS) {
      q->va[id] = qtag(p / QC_SPOTS);
      *reclaim = c;
    }
    return;
  }
  uint32_t c = *cursor != MARKER ? vl_locate(s, *cursor, p) : MARKER;
  if (c == MARKER) c = vl_locate(s, q->front_qc, p);
  if (c == MARKER) {
    s->error |= ERR_CORRUPT;
    return;
  }
  chunk_t* x = ch(s, c);
  x->slots[p % QC_SPOTS] = TOMB;
  const uint32_t old = x->qc_count;
  x->qc_count -= 0x10000u;
  if (old - 0x10000u == QC_SPOTS) *reclaim = vl_set_front(s, q, c) | 0x80000000u;
}

static void finish_reclaim(ouro_sim* s, inst_t* I, cls_t* q, uint32_t reclaim);

/* Remove an exhausted chunk wherever it sits, at the position its queue_pos names (the device's
 * cq_remove). The caller compacts the front once for the wave. Returns 1 when removed. */
static int cq_remove(ouro_sim* s, inst_t* I, cls_t* q, uint32_t chunk, uint32_t* cursor, uint32_t* reclaim) {
  chunk_t* x = ch(s, chunk);
  if (x->queue_pos == NOT_QUEUED || x->queue_pos == QUEUE_PENDING || x->queue_pos == RECYCLED) return 0;
  const uint32_t p = (uint32_t)x->queue_pos;
  x->queue_pos = NOT_QUEUED;
  uint32_t rc = MARKER;
  tomb_slot(s, I, q, p, cursor, &rc);
  if (rc != MARKER) {
    if (*reclaim != MARKER) finish_reclaim(s, I, q, *reclaim);
    *reclaim = rc;
  }
  q->count--;
  return 1; /* the wave compacts front_ once afterwards (inst_malloc) */
}

/* ------------------------------------------------------------------ carves */
static uint64_t capacity(int kind, const cls_t* q) {
  if (is_vl(kind)) return 0xFFFFFFFFull;
  if (is_va(kind)) return (uint64_t)q->ring_size * QC_SPOTS - 2ull * QC_SPOTS;
  return q->ring_size;
}

/* Enqueue chunk indices in lane order (ChunkQueue::enqueue ChunkQueue_impl.cuh:20-38). */
static void enqueue_chunks(ouro_sim* s, inst_t* I, cls_t** qs, int n, const uint32_t* chunks) {
  uint32_t v[64];
  for (int i = 0; i < n; ++i) {
    qs[i]->count++;
    v[i] = qs[i]->back++;
    ch(s, chunks[i])->queue_pos = v[i];
  }
  put_lanes(s, I, qs, n, v, chunks);
}

/* PageQueue allocation lambda + enqueueChunk (PageQueue_impl.cuh:108-133, :58-79), k chunks. */
static uint32_t carve(ouro_sim* s, inst_t* I, cls_t* q, uint32_t k) {
  uint32_t chunks[64];
  uint32_t ok = alloc_chunks(s, &I->reuse, I->add, k, chunks, 0);
  const uint32_t unit = is_chunk_kind(I->kind) ? 1u : q->ppc;
  if (ok) {
    if ((uint64_t)q->carved + (uint64_t)ok * unit > capacity(I->kind, q)) {
      s->error |= ERR_QUEUE_FULL;
      ok = 0;
    } else {
      q->carved += ok * unit;
    }
  }
  for (uint32_t j = 0; j < ok; ++j) {
    chunk_t* x = ch(s, chunks[j]);
    x->page_size = q->page_size;
    x->size = q->ppc;
    if (is_chunk_kind(I->kind)) {
      /* ChunkIndexChunk with all pages free (ChunkAccess.cuh:69-89), 64-bit words */
      x->type = 2;
      x->count = (int32_t)q->ppc;
      x->queue_pos = NOT_QUEUED;
      for (uint32_t w = 0; w < 8; ++w) {
        const uint32_t lo = w * 64;
        x->mask[w] = q->ppc > lo ? ((q->ppc - lo >= 64) ? ~0ull : ((1ull << (q->ppc - lo)) - 1ull)) : 0ull;
      }
    } else {
      x->type = 1; /* PageChunk (PageIndexChunk.cuh:74-78) */
    }
  }
  if (is_chunk_kind(I->kind)) {
    cls_t* qs[64];
    for (uint32_t j = 0; j < ok; ++j) qs[j] = q;
    enqueue_chunks(s, I, qs, (int)ok, chunks);
  } else {
    /* carved pages, then one null entry per page of a chunk the pool could not supply */
    const uint32_t n = k * q->ppc, n_real = ok * q->ppc;
    const uint32_t v0 = q->back;
    q->back += n;
    /* the device writes in batches of up to 64 positions; queue-chunk allocation order is the
       position order either way */
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
      cls_t* qs[64];
      uint32_t v[64], val[64];
      int m = 0;
      for (uint32_t e = e0; e < n && e < e0 + 64; ++e, ++m) {
        qs[m] = q;
        v[m] = v0 + e;
        val[m] = e < n_real ? make_index(chunks[e / q->ppc], e % q->ppc) : NULLPAGE;
      }
      put_lanes(s, I, qs, m, v, val);
    }
  }
  /* signalExpected (BulkSemaphore_impl.cuh:305-308) after publication; page queues publish every
     promised page (nulls included), chunk queues only the carved ones */
  q->cnt += (int64_t)(is_chunk_kind(I->kind) ? ok : k) * q->ppc;
  q->exp -= (int64_t)k * q->ppc;
  return ok;
}

/* ------------------------------------------------------------------ wave ops */
typedef struct {
  int lanes[MAXL];
  int n;
  uint32_t qi;
  int removed; /* the group's claim removed a chunk from its queue */
} group_t;

/* Groups of lanes with equal class, ordered by their lowest lane (match<64> + peeling). */
static int make_groups(const int* lanes, int n, const uint32_t* cls, group_t* g) {
  int ng = 0;
  for (int i = 0; i < n; ++i) {
    int found = -1;
    for (int j = 0; j < ng; ++j)
      if (g[j].qi == cls[i]) found = j;
    if (found < 0) {
      found = ng++;
      g[found].n = 0;
      g[found].qi = cls[i];
      g[found].removed = 0;
    }
    g[found].lanes[g[found].n++] = lanes[i];
  }
  return ng;
}

/* BulkSemaphore::wait (BulkSemaphore_impl.cuh:199-300) for every group leader of one wave: each
 * pass decides for all pending leaders, then carves run in leader order. */
/* Claims in lane order of the group leaders (one fetch-add each on the device, on different words),
 * then the promised carves, one group at a time in lane order. */
static void acquire(ouro_sim* s, inst_t* I, group_t* g, int ng, int* ok, int nact) {
  uint32_t carve_k[64];
  for (int j = 0; j < ng; ++j) {
    cls_t* q = &I->q[g[j].qi];
    const int64_t num = g[j].n;
    carve_k[j] = sim_sem_share(q->cnt, q->exp, num, q->ppc);
    q->cnt -= num;
    q->exp += (int64_t)carve_k[j] * q->ppc;
    ok[j] = 1;
  }
  /* a carve hands one chunk to each active lane: larger promises go in batches of nact */
  for (int j = 0; j < ng; ++j)
    for (uint32_t d = 0; d < carve_k[j]; d += (uint32_t)nact)
      carve(s, I, &I->q[g[j].qi], carve_k[j] - d < (uint32_t)nact ? carve_k[j] - d : (uint32_t)nact);
}

static uint64_t page_offset(const ouro_sim* s, uint32_t c, uint32_t page, uint32_t page_size) {
  (void)s;
  return (uint64_t)c * STRIDE + META + (uint64_t)page * page_size;
}

static uint32_t mask_words(uint32_t ppc) {
  uint32_t w = (ppc + 63) / 64;
  return w == 0 ? 1 : (w > 8 ? 8 : w);
}

/* Start of a claim of n chunk pages (the device's draw_start): the class's page cursor advances by n
 * and its old value, read as a position (cursor / ppc), is the start if it lies in the window
 * [front, back); one past back is folded into the window, one that front has passed restarts at
 * front (fold_start folds the in-group offsets of the per-lane path). A
 * claim that passes REDRAW_AFTER positions without a page draws again for what it still needs, at
 * most MAX_REDRAWS times. */
static uint32_t fold_start(const cls_t* q, uint32_t h) {
  const uint32_t f = q->front, len = q->back - q->front;
  return (len == 0 || len >= 0x80000000u) ? f : (h - f < len ? h : f + (h - f) % len);
}
static uint32_t draw_start(cls_t* q, uint32_t n) {
  const uint64_t p0 = q->scan;
  const uint32_t h = (uint32_t)(p0 / q->ppc), f = q->front, len = q->back - q->front;
  q->scan += (uint64_t)n;
  if (len == 0 || len >= 0x80000000u) return f;
  if (h - f < len) return h;
  if (h - f < 0x80000000u) return f + (h - f) % len; /* ran past back: folded, the cursor keeps its value */
  /* front passed the cursor: it restarts at front (front's cursor value plus this claim) */
  q->scan = (uint64_t)f * q->ppc + n;
  return f;
}

/* Group claim for chunk-based queues: the leader starts at a spread position of the live queue
 * window (per-class page cursor: position = cursor / ppc) and looks at PEEK slots per step for a chunk queued exactly there
 * with free pages; it reserves pages on the chunk count and takes them from the 64-bit mask words,
 * lowest bits first, starting at word (size - count) / 64. The taker of a chunk's last page removes
 * the chunk wherever it sits. (ChunkQueue::allocPage ChunkQueue_impl.cuh:88-151 and
 * ChunkAccess::allocPage ChunkAccess_impl.cuh:38-135, as the device's W = 64 rule.) */
static void claim_group(ouro_sim* s, inst_t* I, group_t* g, uint64_t* out, uint32_t* reclaim) {
  cls_t* q = &I->q[g->qi];
  uint32_t v = draw_start(q, (uint32_t)g->n);
  uint32_t cursor = MARKER, need = (uint32_t)g->n, assigned = 0, wraps = 0, miss = 0, redraws = 0;
  int guard = 0;
  *reclaim = MARKER;
  while (assigned < (uint32_t)g->n) {
    if (++guard > 10000000) {
      s->error |= ERR_TIMEOUT;
      return;
    }
    if (miss >= REDRAW_AFTER && redraws < MAX_REDRAWS) {
      v = draw_start(q, need);
      cursor = MARKER;
      miss = 0;
      ++redraws;
    }
    const uint32_t b = q->back;
    if (v - b < 0x80000000u || v - q->front >= 0x80000000u) { /* at or past back, or behind front */
      v = q->front;
      cursor = MARKER;
      ++wraps;
    }
    uint32_t n = b - v;
    if (n >= 0x80000000u) n = 0;
    if (n > PEEK) n = PEEK;
    if (!is_std(I->kind) && n > QC_SPOTS - v % QC_SPOTS) n = QC_SPOTS - v % QC_SPOTS;
    if ((wraps >= 2 || n == 0) && (s->error & (ERR_OOM | ERR_QUEUE_FULL))) {
      q->cnt += need; /* the pool ran out: the claim goes back, unserved lanes get nullptr */
      return;
    }
    if (n == 0) { /* claimed pages but an empty window: cannot happen sequentially */
      s->error |= ERR_TIMEOUT;
      return;
    }
    uint32_t sl = MARKER, adv = n;
    const uint32_t i = find_chunk(s, I, q, v, n, &cursor, &sl, &adv);
    if (i == n) {
      v += adv;
      miss += adv;
      continue;
    }
    miss = 0;
    chunk_t* x = ch(s, sl);
    const int32_t cnt = x->count;
    const uint32_t t = (uint32_t)cnt < need ? (uint32_t)cnt : need;
    x->count -= (int32_t)t;
    need -= t;
    const uint32_t nw = mask_words(x->size);
    uint32_t w = ((x->size - (uint32_t)cnt) / 64u) % nw;
    if (cnt == (int32_t)t && cq_remove(s, I, q, sl, &cursor, reclaim)) g->removed = 1;
    v += cnt == (int32_t)t ? i + 1u : i;
    uint32_t t_rem = t;
    int wguard = 0;
    while (t_rem > 0 && wguard++ < 64) {
      const uint64_t mw = x->mask[w];
      uint64_t take = 0;
      uint32_t c = 0;
      for (uint32_t bb = 0; bb < 64 && c < t_rem; ++bb)
        if ((mw >> bb) & 1ull) {
          take |= 1ull << bb;
          ++c;
        }
      if (take == 0) {
        w = (w + 1) % nw;
        continue;
      }
      x->mask[w] &= ~take;
      for (uint32_t bb = 0; bb < 64; ++bb)
        if ((take >> bb) & 1ull) {
          const int lane = g->lanes[assigned++];
          out[lane] = page_offset(s, sl, w * 64 + bb, q->page_size);
        }
      t_rem -= c;
      if (t_rem > 0) w = (w + 1) % nw;
    }
    if (t_rem) s->error |= ERR_TIMEOUT;
  }
}

static void finish_reclaim(ouro_sim* s, inst_t* I, cls_t* q, uint32_t reclaim) {
  if (reclaim == MARKER) return;
  if (is_va(I->kind)) ring_put(s, &s->inst[0].reuse, reclaim);
  else if (is_vl(I->kind)) vl_set_old(s, q, reclaim & 0x7FFFFFFFu);
}

/* Front relocation (the device's relocate_front): when the window is longer than twice the class's
 * queued chunks plus RELOC_SLACK, the live chunk that blocks front is removed where it sits and, if it
 * still has free pages, queued again at back. */
static int relocate_front(ouro_sim* s, inst_t* I, cls_t* q) {
  const uint32_t f = q->front, len = q->back - q->front;
  const int32_t live = q->count;
  if (len >= 0x80000000u || len <= 2u * (uint32_t)(live > 0 ? live : 0) + RELOC_SLACK) return 0;
  uint32_t e = MARKER;
  if (is_std(I->kind)) {
    e = *std_slot(q, f);
  } else {
    uint32_t cur = MARKER, tmp[PEEK];
    peek_slots(s, I, q, f, 1, &cur, tmp, NULL);
    e = tmp[0];
  }
  if (!chunk_ok(s, e)) return 0;
  chunk_t* x = ch(s, e);
  if (x->queue_pos != (uint64_t)f || x->page_size != q->page_size) return 0;
  uint32_t cursor = MARKER, reclaim = MARKER;
  if (!cq_remove(s, I, q, e, &cursor, &reclaim)) return 0;
  finish_reclaim(s, I, q, reclaim);
  if (x->count > 0) {
    x->queue_pos = QUEUE_PENDING;
    cls_t* qs[1] = {q};
    enqueue_chunks(s, I, qs, 1, &e);
  }
  return 1;
}

/* Compaction of a chunk queue's front (the per-lane form, wave = 0, or the wave-wide form, which
 * also relocates a live chunk blocking a too-long window, at most MAX_RELOCATE times). */
static void compact(ouro_sim* s, inst_t* I, cls_t* q, int wave) {
  int more = compact_pass(s, I, q, wave);
  if (!wave) return;
  for (uint32_t r = 0; r < MAX_RELOCATE; ++r) {
    if (!more && !relocate_front(s, I, q)) break;
    more = compact_pass(s, I, q, wave);
  }
}

/* Block-wide claim of one class by a whole wave (the device's block_claim_chunks): up to CLAIM_STEP
 * queue positions at a time (one per lane); a prefix sum over the chunks' free counts
 * divides the pages in position order; chunks emptied are removed at the position they were found
 * (slots tombed in position order); each chunk's pages come from its mask words, lowest bits first;
 * front_ is compacted once, after the claim, if it removed anything. */
#define CLAIM_STEP 64u
static void claim_wave(ouro_sim* s, inst_t* I, group_t* g, uint64_t* out) {
  cls_t* q = &I->q[g->qi];
  const uint32_t total = (uint32_t)g->n;
  uint32_t v = draw_start(q, total);
  uint32_t cursor = MARKER, assigned = 0, wraps = 0, rec = MARKER;
  int guard = 0, any_removed = 0;
  while (assigned < total) { /* (block claims draw no new starts) */
    if (++guard > 10000000) {
      s->error |= ERR_TIMEOUT;
      return;
    }
    const uint32_t b = q->back;
    if (v - b < 0x80000000u || v - q->front >= 0x80000000u) {
      v = q->front;
      cursor = MARKER;
      ++wraps;
    }
    /* CLAIM_STEP positions at a time (a single leader peeks PEEK), inside one queue chunk for VA / VL */
    uint32_t nb = b - v;
    if (nb >= 0x80000000u) nb = 0;
    if (nb > CLAIM_STEP) nb = CLAIM_STEP;
    if (!is_std(I->kind) && nb > QC_SPOTS - v % QC_SPOTS) nb = QC_SPOTS - v % QC_SPOTS;
    if ((wraps >= 2 || nb == 0) && (s->error & (ERR_OOM | ERR_QUEUE_FULL))) {
      q->cnt += total - assigned;
      break;
    }
    if (nb == 0) {
      s->error |= ERR_TIMEOUT;
      return;
    }
    uint32_t sl[CLAIM_STEP];
    for (uint32_t i = 0; i < CLAIM_STEP; ++i) sl[i] = MARKER;
    if (is_std(I->kind)) {
      for (uint32_t i = 0; i < nb; ++i) sl[i] = *std_slot(q, v + i);
    } else {
      uint32_t tmp[PEEK], dead = 0;
      for (uint32_t i0 = 0; i0 < nb; i0 += PEEK) {
        const uint32_t m = nb - i0 < PEEK ? nb - i0 : PEEK;
        peek_slots(s, I, q, v + i0, m, &cursor, tmp, i0 == 0 ? &dead : NULL);
        if (dead) break;
        for (uint32_t i = 0; i < m; ++i) sl[i0 + i] = tmp[i];
      }
      if (dead) { /* the queue chunk holds no live entry: pass the rest of it in one step */
        v += dead;
        continue;
      }
    }
    int32_t cnt[CLAIM_STEP];
    uint32_t t[CLAIM_STEP];
    uint32_t before = 0, took = 0;
    const uint32_t rem = total - assigned;
    for (uint32_t i = 0; i < CLAIM_STEP; ++i) {
      cnt[i] = 0;
      t[i] = 0;
      if (i < nb && chunk_ok(s, sl[i])) {
        const chunk_t* x = ch(s, sl[i]);
        if (x->count > 0 && x->queue_pos == (uint64_t)(v + i) && x->page_size == q->page_size) cnt[i] = x->count;
      }
      if (cnt[i] > 0 && before < rem) t[i] = (uint32_t)cnt[i] < rem - before ? (uint32_t)cnt[i] : rem - before;
      before += cnt[i] > 0 ? (uint32_t)cnt[i] : 0u;
    }
    uint32_t r0[CLAIM_STEP];
    for (uint32_t i = 0; i < CLAIM_STEP; ++i) {
      r0[i] = assigned + took;
      took += t[i];
      if (t[i]) ch(s, sl[i])->count -= (int32_t)t[i];
    }
    /* removals: the batch's emptied chunks are tombed (position order); a queue chunk the batch
     * empties is reclaimed once, held as the claim's pending reclaim */
    for (uint32_t i = 0; i < CLAIM_STEP; ++i) {
      if (!t[i] || t[i] != (uint32_t)cnt[i]) continue;
      chunk_t* x = ch(s, sl[i]);
      if (x->queue_pos != (uint64_t)(v + i)) continue;
      x->queue_pos = NOT_QUEUED;
      uint32_t cur = cursor, rc = MARKER;
      tomb_slot(s, I, q, v + i, &cur, &rc);
      if (rc != MARKER) {
        if (rec != MARKER) finish_reclaim(s, I, q, rec);
        rec = rc;
      }
      q->count--;
      any_removed = 1;
    }
    /* mask claims */
    for (uint32_t i = 0; i < CLAIM_STEP; ++i) {
      if (!t[i]) continue;
      chunk_t* x = ch(s, sl[i]);
      const uint32_t nw = mask_words(x->size);
      uint32_t w = ((x->size - (uint32_t)cnt[i]) / 64u) % nw, left = t[i], r = r0[i];
      int wguard = 0;
      while (left > 0 && wguard++ < 64) {
        const uint64_t mw = x->mask[w];
        uint64_t take = 0;
        uint32_t c = 0;
        for (uint32_t bb = 0; bb < 64 && c < left; ++bb)
          if ((mw >> bb) & 1ull) {
            take |= 1ull << bb;
            ++c;
          }
        if (take == 0) {
          w = (w + 1) % nw;
          continue;
        }
        x->mask[w] &= ~take;
        for (uint32_t bb = 0; bb < 64; ++bb)
          if ((take >> bb) & 1ull) out[g->lanes[r++]] = page_offset(s, sl[i], w * 64 + bb, q->page_size);
        left -= c;
        if (left > 0) w = (w + 1) % nw;
      }
      if (left) s->error |= ERR_TIMEOUT;
    }
    assigned += took;
    v += nb;
  }
  if (any_removed) compact(s, I, q, 1);
  if (rec != MARKER) finish_reclaim(s, I, q, rec);
}

/* The per-lane claim of width 1 (the device's W = 1 path of alloc_chunk_pages: the reference's
 * per-thread ChunkQueue::allocPage, ChunkQueue_impl.cuh:88-151, with ChunkAccess::allocPage,
 * ChunkAccess_impl.cuh:38-135): the lane starts at (p0 + rank) / ppc folded into the window, looks at
 * PEEK positions per step for the first chunk queued there with free pages, takes one page off its
 * count (its last page: the lane removes the chunk where it found it) and claims a bit of the mask:
 * from word (tid + 1) mod words, the first free bit at or after bit tid mod 64, cyclically (the
 * reference starts at word (threadIdx.x + 1) mod 16 of its 32-bit words and rotates by threadIdx.x +
 * blockIdx.x; here words are 64 bits and the wave is its own block). Restated for a wave whose lanes
 * request pairwise different classes: lanes of one class race on the chunk count in an order the
 * hardware does not define. */
static uint32_t fold_start(const cls_t* q, uint32_t h);
static uint32_t draw_start(cls_t* q, uint32_t n);
static int claim_lane_w1(ouro_sim* s, inst_t* I, cls_t* q, uint32_t tid, uint32_t h, uint64_t* out,
                         uint32_t* reclaim) {
  uint32_t v = fold_start(q, h), cursor = MARKER, wraps = 0, miss = 0, redraws = 0;
  int guard = 0;
  for (;;) {
    if (++guard > 10000000) {
      s->error |= ERR_TIMEOUT;
      return 0;
    }
    if (miss >= REDRAW_AFTER && redraws < MAX_REDRAWS) {
      v = draw_start(q, 1);
      cursor = MARKER;
      miss = 0;
      ++redraws;
    }
    const uint32_t b = q->back;
    if (v - b < 0x80000000u || v - q->front >= 0x80000000u) {
      v = q->front;
      cursor = MARKER;
      ++wraps;
    }
    uint32_t n = b - v;
    if (n >= 0x80000000u) n = 0;
    if (n > PEEK) n = PEEK;
    if (!is_std(I->kind) && n > QC_SPOTS - v % QC_SPOTS) n = QC_SPOTS - v % QC_SPOTS;
    if ((wraps >= 2 || n == 0) && (s->error & (ERR_OOM | ERR_QUEUE_FULL))) {
      q->cnt += 1; /* the pool ran out: the claim goes back, the lane gets nullptr */
      return 0;
    }
    if (n == 0) {
      s->error |= ERR_TIMEOUT;
      return 0;
    }
    uint32_t c = MARKER, adv = n;
    const uint32_t i = find_chunk(s, I, q, v, n, &cursor, &c, &adv);
    if (i < n) {
      v += i;
      chunk_t* x = ch(s, c);
      const int32_t old = x->count--;
      const uint32_t nw = mask_words(x->size), offset = tid % 64u;
      uint32_t wi = tid, page = MARKER;
      for (int wg = 0; wg < 16 && page == MARKER; ++wg) {
        ++wi;
        const uint64_t cm = x->mask[wi % nw];
        for (uint32_t j = 0; j < 64; ++j) {
          const uint32_t bit = (j + offset) % 64u;
          if ((cm >> bit) & 1ull) {
            x->mask[wi % nw] &= ~(1ull << bit);
            page = (wi % nw) * 64u + bit;
            break;
          }
        }
      }
      if (page == MARKER) {
        s->error |= ERR_TIMEOUT;
        return 0;
      }
      *out = page_offset(s, c, page, q->page_size);
      return old == 1 && cq_remove(s, I, q, c, &cursor, reclaim);
    }
    v += adv;
    miss += adv;
  }
}

/* nact: active lanes of the whole wave (the device enters each instance's block with all of them). */
static void inst_malloc(ouro_sim* s, inst_t* I, const int* lanes, int n, const uint32_t* sizes, uint64_t* out,
                        int nact) {
  uint32_t cls[64];
  for (int i = 0; i < n; ++i) cls[i] = (uint32_t)sim_class_of(sizes[lanes[i]], I->smallest);
  group_t g[64];
  const int ng = make_groups(lanes, n, cls, g);
  int ok[64];
  acquire(s, I, g, ng, ok, nact);
  if (!is_chunk_kind(I->kind)) {
    /* dequeue positions: front fetch-add per group, rank order (atomicAggInc) */
    cls_t* qs[64] = {0};
    uint32_t v[64] = {0}, got[64];
    int owner[64], m = 0;
    /* lanes in lane order, each with its group's front position */
    uint32_t base[64];
    for (int j = 0; j < ng; ++j)
      if (ok[j]) {
        base[j] = I->q[g[j].qi].front;
        I->q[g[j].qi].front += (uint32_t)g[j].n;
      }
    for (int i = 0; i < n; ++i) {
      const int lane = lanes[i];
      for (int j = 0; j < ng; ++j) {
        if (!ok[j]) continue;
        for (int r = 0; r < g[j].n; ++r)
          if (g[j].lanes[r] == lane) {
            qs[m] = &I->q[g[j].qi];
            v[m] = base[j] + (uint32_t)r;
            owner[m] = lane;
            ++m;
          }
      }
    }
    take_lanes(s, I, qs, m, v, got);
    for (int i = 0; i < m; ++i) {
      const uint32_t c = got[i] >> PAGE_BITS, p = got[i] & 511u;
      if (got[i] == MARKER || got[i] == NULLPAGE || !chunk_ok(s, c) || p >= qs[i]->ppc) {
        if (got[i] != MARKER && got[i] != NULLPAGE) s->error |= ERR_CORRUPT;
        continue;
      }
      out[owner[i]] = page_offset(s, c, p, qs[i]->page_size);
    }
    return;
  }
  uint32_t rec[64];
  for (int j = 0; j < ng; ++j) {
    rec[j] = MARKER;
    if (!ok[j]) continue;
    if (s->agg != 1) {
      claim_group(s, I, &g[j], out, &rec[j]);
      continue;
    }
    /* width 1: the group's cursor draw (one fetch-add, restarted at front when front passed it),
       then each lane on its own; a lane's pending queue-chunk reclaim is finished right away (with
       pairwise different classes the order of lanes does not matter) */
    cls_t* q = &I->q[g[j].qi];
    const uint64_t p0 = q->scan;
    const uint32_t h0 = (uint32_t)(p0 / q->ppc), f = q->front, len = q->back - q->front;
    q->scan += (uint64_t)g[j].n;
    uint64_t base = p0;
    if (len != 0 && len < 0x80000000u && h0 - f >= len && h0 - f >= 0x80000000u) {
      base = (uint64_t)f * q->ppc;
      q->scan = base + (uint64_t)g[j].n;
    }
    for (int r = 0; r < g[j].n; ++r) {
      const int lane = g[j].lanes[r];
      uint32_t rc = MARKER;
      if (claim_lane_w1(s, I, q, (uint32_t)lane, (uint32_t)((base + (uint64_t)r) / q->ppc), &out[lane], &rc))
        g[j].removed = 1;
      if (rc != MARKER) {
        if (rec[j] != MARKER) finish_reclaim(s, I, q, rec[j]);
        rec[j] = rc;
      }
    }
  }
  /* one wave-wide compaction per class queue the claims removed chunks from */
  for (int j = 0; j < ng; ++j) {
    int first = g[j].removed;
    for (int i = 0; i < j; ++i)
      if (g[i].qi == g[j].qi && g[i].removed) first = 0;
    if (first) compact(s, I, &I->q[g[j].qi], 1);
  }
  /* deferred queue-chunk reclamation, by leader lane */
  for (int j = 0; j < ng; ++j) finish_reclaim(s, I, &I->q[g[j].qi], rec[j]);
}

/* Block-wide allocation of n <= MAXL requests by a 64-thread workgroup (Ouroboros::malloc_tile, the
 * request of lane l in sub-tile t at index 64 t + l). Phase 1, instance by instance: class leaders are
 * the lanes of wave 0 indexed by class, so claims and carves run in class order; a chunk-queue class
 * is claimed whole (pages assigned in rank order) before the next; a page-queue class takes one
 * front_ fetch-add for all its requests, ranked by request index. Phase 2 (after the block barrier),
 * sub-tile by sub-tile and instance by instance: the page-queue slots are dequeued. */
static void block_malloc(ouro_sim* s, int n, const uint32_t* sizes, uint64_t* out) {
  enum { NI = 4 };
  group_t g[NI][16];
  int ng[NI] = {0}, ok[NI][16];
  uint32_t base[NI][16];
  for (int i = 0; i < n; ++i) out[i] = UINT64_MAX;
  int lanes[MAXL];
  for (int k = 0; k < s->ninst && k < NI; ++k) {
    inst_t* I = &s->inst[k];
    int nh = 0;
    uint32_t cls[MAXL];
    for (int i = 0; i < n; ++i) {
      if (!sizes[i] || sizes[i] > I->largest) continue;
      int prior = 0;
      for (int j = 0; j < k; ++j) prior |= sizes[i] <= s->inst[j].largest;
      if (prior) continue;
      lanes[nh] = i;
      cls[nh++] = (uint32_t)sim_class_of(sizes[i], I->smallest);
    }
    if (!nh) continue;
    ng[k] = make_groups(lanes, nh, cls, g[k]);
    for (int a = 1; a < ng[k]; ++a)
      for (int b = a; b > 0 && g[k][b - 1].qi > g[k][b].qi; --b) {
        const group_t t = g[k][b];
        g[k][b] = g[k][b - 1];
        g[k][b - 1] = t;
      }
    if (is_chunk_kind(I->kind)) {
      for (int j = 0; j < ng[k]; ++j) {
        acquire(s, I, &g[k][j], 1, &ok[k][j], 64);
        claim_wave(s, I, &g[k][j], out);
      }
    } else {
      acquire(s, I, g[k], ng[k], ok[k], 64);
      for (int j = 0; j < ng[k]; ++j)
        if (ok[k][j]) {
          base[k][j] = I->q[g[k][j].qi].front;
          I->q[g[k][j].qi].front += (uint32_t)g[k][j].n;
        }
    }
  }
  for (int i = 0; i < n; ++i) {
    int fits = 0;
    for (int k = 0; k < s->ninst; ++k) fits |= sizes[i] <= s->inst[k].largest;
    if (sizes[i] && !fits) s->error |= ERR_TOO_LARGE;
  }
  for (int t = 0; t * 64 < n; ++t)
    for (int k = 0; k < s->ninst && k < NI; ++k) {
      inst_t* I = &s->inst[k];
      if (is_chunk_kind(I->kind) || !ng[k]) continue;
      cls_t* qs[64];
      uint32_t v[64], got[64];
      int owner[64], m = 0;
      for (int l = t * 64; l < n && l < t * 64 + 64; ++l)
        for (int j = 0; j < ng[k]; ++j) {
          if (!ok[k][j]) continue;
          for (int r = 0; r < g[k][j].n; ++r)
            if (g[k][j].lanes[r] == l) {
              qs[m] = &I->q[g[k][j].qi];
              v[m] = base[k][j] + (uint32_t)r;
              owner[m++] = l;
            }
        }
      if (!m) continue;
      take_lanes(s, I, qs, m, v, got);
      for (int i = 0; i < m; ++i) {
        const uint32_t c = got[i] >> PAGE_BITS, p = got[i] & 511u;
        if (got[i] == MARKER || got[i] == NULLPAGE || !chunk_ok(s, c) || p >= qs[i]->ppc) {
          if (got[i] != MARKER && got[i] != NULLPAGE) s->error |= ERR_CORRUPT;
          continue;
        }
        out[owner[i]] = page_offset(s, c, p, qs[i]->page_size);
      }
    }
}

void sim_wave_malloc(ouro_sim* s, int n, const uint32_t* sizes, uint64_t* out) {
  if (s->block_order) {
    block_malloc(s, n, sizes, out);
    return;
  }
  int rest[64], nr = 0;
  for (int i = 0; i < n; ++i) {
    out[i] = UINT64_MAX;
    if (sizes[i]) rest[nr++] = i;
  }
  const int nact = nr;
  for (int k = 0; k < s->ninst && nr; ++k) {
    inst_t* I = &s->inst[k];
    int here[64], nh = 0, later[64], nl = 0;
    for (int i = 0; i < nr; ++i) {
      if (sizes[rest[i]] <= I->largest) here[nh++] = rest[i];
      else later[nl++] = rest[i];
    }
    if (nh) inst_malloc(s, I, here, nh, sizes, out, nact);
    memcpy(rest, later, sizeof(int) * (size_t)nl);
    nr = nl;
  }
  if (nr) s->error |= ERR_TOO_LARGE;
}

/* ---- free */
/* Chunk recycling (the device's flash_chunks; ChunkQueue::freePage's flash, ChunkQueue_impl.cuh:163-199).
 * Groups whose free made their chunk full, in leader-lane order and before any enqueue or signal of
 * this call: when the class queue holds more than recycle_min chunks (a chunk that the free found
 * exhausted, so out of the queue, counting itself) and the semaphore has the chunk's other
 * size - n pages unclaimed, they leave the semaphore, the chunk leaves its queue (tombstone, count_)
 * and its class (carved, page size 0). Then, as the wave does: the flashed chunks go to the
 * instance's reuse ring in that order (those the ring has no room for return to their class with
 * every page free and are queued again), the queues that lost entries are compacted (wave form) in
 * the order of their first flash, and queue chunks emptied by the tombstones are reclaimed.
 * gone[j] marks the groups whose pages left with their chunk. */
static void flash(ouro_sim* s, inst_t* I, const int* lanes, int n, const uint32_t* qi, const group_t* g, int ng,
                  const int* unqueued, int* gone) {
  cls_t* rq[64];
  cls_t* fq[64];
  cls_t* recq[64];
  uint32_t rec[64], fc[64];
  int64_t fred[64];
  int fj[64];
  int nrq = 0, nrec = 0, nf = 0;
  for (int j = 0; j < ng; ++j) {
    const uint32_t c = g[j].qi;
    chunk_t* x = ch(s, c);
    int idx = 0;
    for (int i = 0; i < n; ++i)
      if (lanes[i] == g[j].lanes[0]) idx = i;
    cls_t* q = &I->q[qi[idx]];
    if (x->count != (int32_t)x->size || (int64_t)q->count + unqueued[j] <= (int64_t)s->recycle_min) continue;
    const int64_t reduce = (int64_t)x->size - g[j].n;
    if (q->cnt < reduce) continue;
    q->cnt -= reduce;
    x->count = 0;
    if (x->queue_pos < 0x100000000ull) {
      uint32_t cursor = MARKER, rc = MARKER;
      tomb_slot(s, I, q, (uint32_t)x->queue_pos, &cursor, &rc);
      if (rc != MARKER) {
        recq[nrec] = q;
        rec[nrec++] = rc;
      }
      q->count--;
      int seen = 0;
      for (int r = 0; r < nrq; ++r) seen |= rq[r] == q;
      if (!seen) rq[nrq++] = q;
    } else if (x->queue_pos != NOT_QUEUED) {
      s->error |= ERR_TIMEOUT; /* an enqueue in flight cannot exist in the sequential model */
    }
    x->queue_pos = RECYCLED;
    q->carved--;
    x->page_size = 0;
    fq[nf] = q;
    fc[nf] = c;
    fred[nf] = reduce;
    fj[nf++] = j;
  }
  cls_t* bq[64];
  uint32_t bc[64];
  int nb = 0;
  /* the wave's flashed chunks go into the ring together or not at all (a quarter stays free) */
  const int fits = I->reuse.count + nf <= (int32_t)(I->reuse.size - I->reuse.size / 4);
  for (int f = 0; f < nf; ++f) {
    chunk_t* x = ch(s, fc[f]);
    if (fits) {
      x->type = 0;
      ring_put(s, &I->reuse, fc[f]);
      gone[fj[f]] = 1;
    } else {
      /* no room in the ring: back to its class, every page free */
      x->page_size = fq[f]->page_size;
      fq[f]->carved++;
      x->count = (int32_t)x->size;
      x->queue_pos = QUEUE_PENDING;
      bq[nb] = fq[f];
      bc[nb++] = fc[f];
    }
  }
  if (nb) enqueue_chunks(s, I, bq, nb, bc);
  for (int f = 0; f < nf; ++f)
    if (!gone[fj[f]]) fq[f]->cnt += fred[f];
  for (int r = 0; r < nrq; ++r) compact(s, I, rq[r], 1);
  for (int r = 0; r < nrec; ++r) finish_reclaim(s, I, recq[r], rec[r]);
}

static void inst_free(ouro_sim* s, inst_t* I, const int* lanes, int n, const uint32_t* qi, const uint32_t* chunk,
                      const uint32_t* page) {
  if (!is_chunk_kind(I->kind)) {
    /* back positions per class group in lane order, publish, then signal (PageQueue_impl.cuh:137-143) */
    cls_t* qs[64] = {0};
    uint32_t v[64] = {0}, val[64] = {0};
    uint32_t cls[64] = {0};
    for (int i = 0; i < n; ++i) cls[i] = qi[i];
    group_t g[64];
    const int ng = make_groups(lanes, n, cls, g);
    uint32_t base[64];
    for (int j = 0; j < ng; ++j) {
      base[j] = I->q[g[j].qi].back;
      I->q[g[j].qi].back += (uint32_t)g[j].n;
    }
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < ng; ++j)
        for (int r = 0; r < g[j].n; ++r)
          if (g[j].lanes[r] == lanes[i]) {
            qs[i] = &I->q[g[j].qi];
            v[i] = base[j] + (uint32_t)r;
          }
      val[i] = make_index(chunk[i], page[i]);
    }
    put_lanes(s, I, qs, n, v, val);
    for (int i = 0; i < n; ++i) qs[i]->cnt += 1;
    return;
  }
  /* ChunkAccess::freePage (ChunkAccess_impl.cuh:7-25) grouped by chunk */
  for (int i = 0; i < n; ++i) ch(s, chunk[i])->mask[page[i] / 64] |= 1ull << (page[i] % 64);
  group_t g[64];
  const int ng = make_groups(lanes, n, chunk, g); /* groups keyed by chunk index */
  uint32_t ff_chunk[64];
  cls_t* ff_q[64];
  int ff_j[64], unqueued[64] = {0};
  int nff = 0;
  for (int j = 0; j < ng; ++j) {
    const uint32_t c = g[j].qi;
    int idx = 0;
    for (int i = 0; i < n; ++i)
      if (lanes[i] == g[j].lanes[0]) idx = i;
    cls_t* q = &I->q[qi[idx]];
    chunk_t* x = ch(s, c);
    const int32_t old = x->count;
    x->count += g[j].n;
    /* count crossed from <= 0 to > 0: the chunk must be in the queue; NOT_QUEUED means it left */
    unqueued[j] = old <= 0 && old + g[j].n > 0;
    if (unqueued[j] && x->queue_pos == NOT_QUEUED) {
      ff_j[nff] = j;
      ff_chunk[nff] = c;
      ff_q[nff++] = q;
    }
  }
  /* recycling first: a chunk recycled here is not queued again */
  int gone[64] = {0};
  if (s->recycle_min != MARKER) flash(s, I, lanes, n, qi, g, ng, unqueued, gone);
  int m = 0;
  for (int f = 0; f < nff; ++f)
    if (!gone[ff_j[f]]) {
      ch(s, ff_chunk[f])->queue_pos = QUEUE_PENDING;
      ff_chunk[m] = ff_chunk[f];
      ff_q[m++] = ff_q[f];
    }
  if (m) enqueue_chunks(s, I, ff_q, m, ff_chunk);
  for (int j = 0; j < ng; ++j) {
    int idx = 0;
    for (int i = 0; i < n; ++i)
      if (lanes[i] == g[j].lanes[0]) idx = i;
    if (!gone[j]) I->q[qi[idx]].cnt += g[j].n;
  }
}

void sim_wave_free(ouro_sim* s, int n, const uint64_t* offsets) {
  int lanes1[2][64], n1[2] = {0, 0};
  uint32_t qi[2][64], chunk[2][64], page[2][64];
  const uint32_t nwords = (uint32_t)((s->max_chunks + 31) / 32);
  for (int i = 0; i < n; ++i) {
    const uint64_t off = offsets[i];
    if (off == UINT64_MAX) continue;
    if (off >= s->max_chunks * STRIDE) {
      s->error |= ERR_BAD_FREE;
      continue;
    }
    const uint32_t c = sim_locate(s->locator, nwords, (uint32_t)(off / STRIDE));
    if (c == MARKER) {
      s->error |= ERR_BAD_FREE;
      continue;
    }
    const uint32_t ps = ch(s, c)->page_size;
    const uint64_t base = (uint64_t)c * STRIDE + META;
    if (ps == 0 || off < base) {
      s->error |= ERR_BAD_FREE;
      continue;
    }
    int k = 0;
    while (k < s->ninst && ps > s->inst[k].largest) ++k;
    if (k == s->ninst) {
      s->error |= ERR_BAD_FREE;
      continue;
    }
    inst_t* I = &s->inst[k];
    const uint64_t rel = off - base;
    if (ps < I->smallest || (ps & (ps - 1)) || rel % ps || rel / ps >= I->chunk_size / ps) {
      s->error |= ERR_BAD_FREE;
      continue;
    }
    lanes1[k][n1[k]] = i;
    qi[k][n1[k]] = (uint32_t)sim_class_of(ps, I->smallest);
    chunk[k][n1[k]] = c;
    page[k][n1[k]] = (uint32_t)(rel / ps);
    n1[k]++;
  }
  for (int k = 0; k < s->ninst; ++k)
    if (n1[k]) inst_free(s, &s->inst[k], lanes1[k], n1[k], qi[k], chunk[k], page[k]);
}

/* ------------------------------------------------------------------ create / state */
ouro_sim* sim_create(int variant, uint64_t pool_bytes) {
  if (variant < 0 || variant > 5) return NULL;
  ouro_sim* s = (ouro_sim*)calloc(1, sizeof(ouro_sim));
  s->variant = variant;
  s->agg = 64;
  s->recycle_min = MARKER;
  sim_layout(variant, pool_bytes, &s->max_chunks, &s->alloc_size);
  /* a virtualized pool keeps its top chunks for queue chunks (at most a quarter of a tiny pool) */
  s->qc_reserve = is_std(variant) ? 0u : (uint32_t)(QC_RESERVE < s->max_chunks / 4 ? QC_RESERVE : s->max_chunks / 4);
  s->chunks = (chunk_t*)calloc(s->max_chunks + 4, sizeof(chunk_t));
  s->locator = (uint32_t*)calloc((s->max_chunks + 31) / 32 + 1, 4);
  s->ninst = 2;
  uint32_t qc_next = 0;
  for (int k = 0; k < 2; ++k) {
    inst_t* I = &s->inst[k];
    I->kind = variant;
    I->smallest = k == 0 ? 16u : (16u << 10);
    I->nq = k == 0 ? 10u : 2u;
    I->largest = I->smallest << (I->nq - 1);
    I->chunk_size = I->largest;
    I->add = I->chunk_size / BASE_CHUNK;
    I->reuse.size = REUSE_ENTRIES;
    I->reuse.ring = (uint32_t*)malloc(REUSE_ENTRIES * 4);
    memset(I->reuse.ring, 0xFF, REUSE_ENTRIES * 4);
    for (uint32_t i = 0; i < I->nq; ++i) {
      cls_t* q = &I->q[i];
      q->page_size = I->smallest << i;
      q->ppc = I->chunk_size / q->page_size;
      q->ring_size = (uint32_t)entries_of(variant);
      q->front_qc = q->back_qc = q->old_qc = MARKER;
      if (is_std(variant)) q->blocks = (uint32_t**)calloc(q->ring_size / BLOCK, sizeof(uint32_t*));
      if (!is_std(variant)) {
        /* PageQueueVA::init etc. take queue chunks 0, 1, 2, ... (PageQueueVA_impl.cuh:13-28) */
        const uint32_t c = qc_next++;
        qc_init(s, c, 0);
        s->locator[c >> 5] |= 1u << (c & 31u);
        if (is_va(variant)) {
          q->va = (uint32_t*)malloc(q->ring_size * 4);
          memset(q->va, 0xFF, q->ring_size * 4);
          q->va[0] = c;
        } else {
          q->front_qc = q->back_qc = q->old_qc = c;
        }
      }
    }
  }
  s->next_free = qc_next;
  return s;
}

void sim_destroy(ouro_sim* s) {
  if (!s) return;
  for (int k = 0; k < 2; ++k) {
    inst_t* I = &s->inst[k];
    free(I->reuse.ring);
    for (uint32_t i = 0; i < I->nq; ++i) {
      cls_t* q = &I->q[i];
      if (q->blocks) {
        for (uint32_t b = 0; b < q->ring_size / BLOCK; ++b) free(q->blocks[b]);
        free(q->blocks);
      }
      free(q->va);
    }
  }
  for (uint64_t c = 0; c < s->max_chunks + 4; ++c) free(s->chunks[c].slots);
  free(s->chunks);
  free(s->locator);
  free(s);
}

void sim_get_state(const ouro_sim* s, sim_state* out) {
  memset(out, 0, sizeof(*out));
  out->next_free_chunk = s->next_free;
  out->error = s->error;
  uint32_t k = 0;
  for (int i = 0; i < s->ninst; ++i) {
    const inst_t* I = &s->inst[i];
    out->reuse_count[i] = I->reuse.count;
    out->reuse_front[i] = I->reuse.front;
    out->reuse_back[i] = I->reuse.back;
    for (uint32_t j = 0; j < I->nq; ++j, ++k) {
      const cls_t* q = &I->q[j];
      out->sem_count[k] = q->cnt;
      out->sem_expected[k] = (uint32_t)q->exp;
      out->sem_reserved[k] = (uint32_t)q->res;
      out->front[k] = q->front;
      out->back[k] = q->back;
      out->queue_count[k] = q->count;
      out->carved[k] = q->carved;
    }
  }
  out->num_classes = k;
}

int sim_chunk_info(const ouro_sim* s, uint32_t c, uint32_t* page_size, int32_t* count, uint64_t* mask8) {
  if (c >= s->max_chunks) return -1;
  const chunk_t* x = &s->chunks[c];
  if (page_size) *page_size = x->page_size;
  if (count) *count = x->count;
  if (mask8) memcpy(mask8, x->mask, sizeof(x->mask));
  return x->type;
}

uint64_t sim_chunk_queue_pos(const ouro_sim* s, uint32_t c) {
  return c < s->max_chunks ? s->chunks[c].queue_pos : NOT_QUEUED;
}

/* Slot values of class k's queue window [front, back) (k counts the classes of both instances), up to
 * max entries; returns the window length. Reads only: dead VA / VL queue chunks read as TOMB, positions
 * whose queue chunk is missing as MARKER. */
uint32_t sim_queue_window(const ouro_sim* s, uint32_t k, uint32_t* out, uint32_t max) {
  const inst_t* I = &s->inst[0];
  if (k >= I->nq) {
    k -= I->nq;
    I = &s->inst[1];
    if (k >= I->nq) return 0;
  }
  const cls_t* q = &I->q[k];
  const uint32_t len = q->back - q->front;
  for (uint32_t i = 0; i < len && i < max; ++i) {
    const uint32_t v = q->front + i;
    uint32_t val = MARKER;
    if (is_std(I->kind)) {
      const uint32_t p = v & (q->ring_size - 1);
      const uint32_t* b = q->blocks[p / BLOCK];
      val = b ? b[p % BLOCK] : MARKER;
    } else if (is_va(I->kind)) {
      const uint32_t e = q->va[(v / QC_SPOTS) & (q->ring_size - 1)];
      if (e == qtag(v / QC_SPOTS)) val = TOMB;
      else if (chunk_ok(s, e) && s->chunks[e].slots && s->chunks[e].vstart / QC_SPOTS == v / QC_SPOTS)
        val = s->chunks[e].slots[v % QC_SPOTS];
    } else {
      uint32_t c = q->front_qc;
      const uint32_t vb = v - v % QC_SPOTS;
      if (chunk_ok(s, c) && s->chunks[c].vstart - vb < 0x80000000u && s->chunks[c].vstart != vb) {
        val = TOMB;
      } else {
        while (chunk_ok(s, c) && s->chunks[c].vstart != vb) c = s->chunks[c].next;
        if (chunk_ok(s, c)) val = s->chunks[c].slots[v % QC_SPOTS];
      }
    }
    out[i] = val;
  }
  return len;
}

void sim_set_block_order(ouro_sim* s, int on) { s->block_order = on; }
/* enqueueInitialChunk on a chunk taken from the pool (the device's preloadChunk; PageQueue_impl.cuh:40-53,
 * ChunkQueue_impl.cuh:72-78): pages [ppc - available, ppc) become free pages of chain class k, the
 * others stay with the caller. Returns the chunk, or MARKER when the pool has none. */
int sim_set_recycling(ouro_sim* s, uint32_t min_queued) {
  if (!is_chunk_kind(s->inst[0].kind)) return -1;
  s->recycle_min = min_queued;
  return 0;
}

uint32_t sim_preload(ouro_sim* s, uint32_t k, uint32_t available) {
  int i = 0;
  while (i < s->ninst && k >= s->inst[i].nq) k -= s->inst[i++].nq;
  if (i == s->ninst) return MARKER;
  inst_t* I = &s->inst[i];
  cls_t* q = &I->q[k];
  uint32_t chunk = MARKER;
  if (alloc_chunks(s, &I->reuse, I->add, 1, &chunk, 0) != 1) return MARKER;
  const uint32_t avail = available > q->ppc ? q->ppc : available, taken = q->ppc - avail;
  chunk_t* x = ch(s, chunk);
  x->page_size = q->page_size;
  x->size = q->ppc;
  if (is_chunk_kind(I->kind)) {
    x->type = 2;
    x->count = (int32_t)avail;
    x->queue_pos = NOT_QUEUED;
    for (uint32_t w = 0; w < 8; ++w) {
      x->mask[w] = 0;
      for (uint32_t b = 0; b < 64; ++b)
        if (w * 64 + b >= taken && w * 64 + b < q->ppc) x->mask[w] |= 1ull << b;
    }
    q->carved += 1;
    if (avail) {
      cls_t* qs[1] = {q};
      enqueue_chunks(s, I, qs, 1, &chunk);
    }
  } else {
    x->type = 1;
    q->carved += q->ppc;
    const uint32_t v0 = q->back;
    q->back += avail;
    for (uint32_t e0 = 0; e0 < avail; e0 += 64) {
      cls_t* qs[64];
      uint32_t v[64], val[64];
      int m = 0;
      for (uint32_t e = e0; e < avail && e < e0 + 64; ++e, ++m) {
        qs[m] = q;
        v[m] = v0 + e;
        val[m] = make_index(chunk, taken + e);
      }
      put_lanes(s, I, qs, m, v, val);
    }
  }
  q->cnt += avail;
  return chunk;
}

int sim_set_agg(ouro_sim* s, int w) {
  if (w != 1 && w != 64) return -1;
  s->agg = w;
  return 0;
}

/* Start every class queue at position `origin` (a multiple of QC_SPOTS) on a fresh pool: front, back,
 * the page cursor, and for VA / VL the initial queue chunk's virtual start (VA: its pointer-ring
 * entry moves to the entry of origin's queue chunk). Positions are 32-bit and wrap; a trace from any
 * origin hands out the same blocks as from 0. */
int sim_set_origin(ouro_sim* s, uint32_t origin) {
  if (origin % QC_SPOTS) return -1;
  for (int k = 0; k < s->ninst; ++k) {
    inst_t* I = &s->inst[k];
    for (uint32_t i = 0; i < I->nq; ++i) {
      cls_t* q = &I->q[i];
      if (q->front != 0 || q->back != 0 || q->cnt != 0) return -1;
      q->front = q->back = origin;
      q->scan = (uint64_t)origin * q->ppc;
      if (is_va(I->kind)) {
        const uint32_t c = q->va[0];
        q->va[0] = MARKER;
        q->va[(origin / QC_SPOTS) & (q->ring_size - 1)] = c;
        ch(s, c)->vstart = origin;
      } else if (is_vl(I->kind)) {
        ch(s, q->front_qc)->vstart = origin;
      }
    }
  }
  return 0;
}
