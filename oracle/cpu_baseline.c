/*
 * CPU baseline -- TEST / BENCH INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg).
 *
 * The identical request trace served by glibc malloc/free, one POSIX thread per requested core,
 * each owning a contiguous slice of the "device threads". Every round runs malloc -> write the
 * thread id into every 32-bit word -> verify -> free, with barriers between phases; the malloc
 * and free phases are timed separately, as the device harness times its alloc and free kernels
 * (src/main.cu:146-167).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  const uint32_t* sizes;
  void** ptrs;
  uint32_t lo, hi, rounds;
  pthread_barrier_t* bar;
  double* t_alloc;
  double* t_free;
  uint64_t errors;
  int tid;
} arg_t;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void* worker(void* p) {
  arg_t* a = (arg_t*)p;
  for (uint32_t r = 0; r < a->rounds; ++r) {
    pthread_barrier_wait(a->bar);
    double t0 = now_s();
    for (uint32_t i = a->lo; i < a->hi; ++i) a->ptrs[i] = malloc(a->sizes[i]);
    pthread_barrier_wait(a->bar);
    if (a->tid == 0) *a->t_alloc += now_s() - t0;
    for (uint32_t i = a->lo; i < a->hi; ++i) {
      uint32_t* w = (uint32_t*)a->ptrs[i];
      if (!w) { a->errors++; continue; }
      for (uint32_t k = 0; k < a->sizes[i] / 4; ++k) w[k] = i;
    }
    for (uint32_t i = a->lo; i < a->hi; ++i) {
      const uint32_t* w = (const uint32_t*)a->ptrs[i];
      if (!w) continue;
      for (uint32_t k = 0; k < a->sizes[i] / 4; ++k) a->errors += (w[k] != i);
    }
    pthread_barrier_wait(a->bar);
    t0 = now_s();
    for (uint32_t i = a->lo; i < a->hi; ++i) free(a->ptrs[i]);
    pthread_barrier_wait(a->bar);
    if (a->tid == 0) *a->t_free += now_s() - t0;
  }
  return NULL;
}

/* Returns 0 on success. alloc_s / free_s are summed over rounds (wall time of the phase). */
int cpu_glibc_trace(const uint32_t* sizes, uint32_t n, uint32_t rounds, int threads, double* alloc_s,
                    double* free_s, uint64_t* errors) {
  if (threads < 1) threads = 1;
  void** ptrs = (void**)calloc(n, sizeof(void*));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  arg_t* args = (arg_t*)calloc((size_t)threads, sizeof(arg_t));
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)threads);
  double ta = 0, tf = 0;
  for (int t = 0; t < threads; ++t) {
    args[t].sizes = sizes;
    args[t].ptrs = ptrs;
    args[t].lo = (uint32_t)((uint64_t)n * t / threads);
    args[t].hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
    args[t].rounds = rounds;
    args[t].bar = &bar;
    args[t].t_alloc = &ta;
    args[t].t_free = &tf;
    args[t].tid = t;
    pthread_create(&th[t], NULL, worker, &args[t]);
  }
  uint64_t err = 0;
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    err += args[t].errors;
  }
  pthread_barrier_destroy(&bar);
  *alloc_s = ta;
  *free_s = tf;
  *errors = err;
  free(args);
  free(th);
  free(ptrs);
  return 0;
}
