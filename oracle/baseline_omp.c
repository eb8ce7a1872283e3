/*
 * baseline_omp.c -- OpenMP CPU baselines with the oracle's semantics.
 *
 * TEST/MEASUREMENT INFRASTRUCTURE: timed by bench.py's cpu_baseline leg and
 * checked against the serial oracle by tests/.  The reference has no CPU
 * path; these are this project's own parallel restatements of
 * samples/bfs/bfs.cu, samples/sssp/sssp.cu, samples/pagerank/pr.cu and
 * samples/cc/cc.cu.
 *
 *   bl_bfs       direction-optimizing BFS (top-down queue / bottom-up
 *                bitmap, Beamer et al. SC'12 switch with alpha=14, beta=24)
 *   bl_sssp      near-far delta-stepping with CAS atomic-min relaxation
 *   bl_pagerank  FP32 pull, same recurrence as or_pagerank
 *   bl_cc        CAS hooking union-find (root = smaller id) + compression
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

static double now_ms(void) { return omp_get_wtime() * 1e3; }
int bl_max_threads(void) { return omp_get_max_threads(); }
static int nth(int n) { return n > 0 ? n : omp_get_max_threads(); }

/* ------------------------------------------------------------------ */
double bl_bfs(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
              uint32_t *level, int nthreads) {
  int T = nth(nthreads);
  uint64_t W = (V + 63) / 64;
  uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (V + 1));
  uint32_t *nq = (uint32_t *)malloc(sizeof(uint32_t) * (V + 1));
  uint64_t *fb = (uint64_t *)calloc(W + 1, sizeof(uint64_t));
  uint64_t *nb = (uint64_t *)calloc(W + 1, sizeof(uint64_t));
  uint64_t *tcount = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t i = 0; i < V; i++) level[i] = OR_INFINIT;
  double t0 = now_ms();
  if (src >= V) goto done;
  level[src] = 0;
  q[0] = src;
  uint64_t nf = 1, E = off[V];
  uint64_t mf = off[src + 1] - off[src], mu = E - mf;
  int bottom_up = 0;
  uint32_t d = 0;
  while (nf) {
    uint64_t new_nf = 0, new_mf = 0;
    if (!bottom_up && mf > mu / 14) {
      /* queue -> bitmap */
      memset(fb, 0, W * 8);
#pragma omp parallel for num_threads(T) schedule(static)
      for (uint64_t i = 0; i < nf; i++)
        __atomic_fetch_or(&fb[q[i] >> 6], 1ull << (q[i] & 63), __ATOMIC_RELAXED);
      bottom_up = 1;
    } else if (bottom_up && nf < V / 24) {
      /* bitmap -> queue */
      uint64_t n = 0;
      for (uint64_t wi = 0; wi < W; wi++) {
        uint64_t m = fb[wi];
        while (m) {
          int b = __builtin_ctzll(m);
          m &= m - 1;
          q[n++] = (uint32_t)(wi * 64 + b);
        }
      }
      nf = n;
      bottom_up = 0;
    }
    if (bottom_up) {
      memset(nb, 0, W * 8);
#pragma omp parallel for num_threads(T) schedule(dynamic, 64) reduction(+ : new_nf, new_mf)
      for (uint64_t wi = 0; wi < W; wi++) {
        uint64_t found = 0;
        uint64_t vend = (wi + 1) * 64 < V ? (wi + 1) * 64 : V;
        for (uint64_t v = wi * 64; v < vend; v++) {
          if (level[v] != OR_INFINIT) continue;
          for (uint64_t e = off[v]; e < off[v + 1]; e++) {
            uint32_t u = col[e];
            if (fb[u >> 6] >> (u & 63) & 1) {
              level[v] = d + 1;
              found |= 1ull << (v & 63);
              new_nf++;
              new_mf += off[v + 1] - off[v];
              break;
            }
          }
        }
        nb[wi] = found;
      }
      uint64_t *t = fb;
      fb = nb;
      nb = t;
    } else {
      /* top-down: thread-local append then concatenate */
#pragma omp parallel num_threads(T) reduction(+ : new_mf)
      {
        int tid = omp_get_thread_num();
        uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * 4096);
        uint64_t cap = 4096, n = 0;
#pragma omp for schedule(dynamic, 16)
        for (uint64_t i = 0; i < nf; i++) {
          uint32_t u = q[i];
          for (uint64_t e = off[u]; e < off[u + 1]; e++) {
            uint32_t v = col[e];
            if (level[v] != OR_INFINIT) continue;
            uint32_t exp = OR_INFINIT;
            if (__atomic_compare_exchange_n(&level[v], &exp, d + 1, 0, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED)) {
              if (n == cap) {
                cap *= 2;
                buf = (uint32_t *)realloc(buf, sizeof(uint32_t) * cap);
              }
              buf[n++] = v;
              new_mf += off[v + 1] - off[v];
            }
          }
        }
        tcount[tid] = n;
#pragma omp barrier
#pragma omp single
        {
          uint64_t s = 0;
          for (int t = 0; t < T; t++) {
            uint64_t c = tcount[t];
            tcount[t] = s;
            s += c;
          }
          tcount[T] = s;
        }
        memcpy(nq + tcount[tid], buf, n * sizeof(uint32_t));
        free(buf);
      }
      new_nf = tcount[T];
      uint32_t *t = q;
      q = nq;
      nq = t;
    }
    nf = new_nf;
    mu -= new_mf;
    mf = new_mf;
    d++;
  }
done:;
  double t1 = now_ms();
  free(q);
  free(nq);
  free(fb);
  free(nb);
  free(tcount);
  return t1 - t0;
}

/* ------------------------------------------------------------------ */
static inline int atomic_min_u32(uint32_t *p, uint32_t v) {
  uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < old) {
    if (__atomic_compare_exchange_n(p, &old, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return 1;
  }
  return 0;
}

double bl_sssp(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w,
               uint32_t src, uint32_t delta, uint32_t *dist, int nthreads) {
  int T = nth(nthreads);
  if (delta == 0) delta = 1;
  uint32_t *stamp = (uint32_t *)malloc(sizeof(uint32_t) * (V + 1));
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t i = 0; i < V; i++) {
    dist[i] = OR_INFINIT;
    stamp[i] = 0;
  }
  double t0 = now_ms();
  if (src >= V) {
    free(stamp);
    return 0.0;
  }
  uint64_t ncap = 1024, fcap = 1024, nn = 1, nfar = 0;
  uint32_t *near = (uint32_t *)malloc(sizeof(uint32_t) * ncap);
  uint32_t *far = (uint32_t *)malloc(sizeof(uint32_t) * fcap);
  uint64_t *tn = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
  uint64_t *tf = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
  uint32_t **lb_n = (uint32_t **)calloc((size_t)T, sizeof(void *));
  uint32_t **lb_f = (uint32_t **)calloc((size_t)T, sizeof(void *));
  uint64_t *cap_n = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
  uint64_t *cap_f = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
  dist[src] = 0;
  near[0] = src;
  uint64_t thr = delta;
  uint32_t round = 0;
  for (;;) {
    while (nn) {
      round++;
#pragma omp parallel num_threads(T)
      {
        int tid = omp_get_thread_num();
        uint64_t a = 0, b = 0;
#pragma omp for schedule(dynamic, 64)
        for (uint64_t i = 0; i < nn; i++) {
          uint32_t u = near[i];
          uint32_t du = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
          for (uint64_t e = off[u]; e < off[u + 1]; e++) {
            uint32_t v = col[e];
            uint32_t nd = du + (w ? w[e] : 1u);
            if (atomic_min_u32(&dist[v], nd)) {
              if (nd < thr) {
                if (__atomic_exchange_n(&stamp[v], round, __ATOMIC_RELAXED) != round) {
                  if (a == cap_n[tid]) {
                    cap_n[tid] = cap_n[tid] ? cap_n[tid] * 2 : 1024;
                    lb_n[tid] = (uint32_t *)realloc(lb_n[tid], sizeof(uint32_t) * cap_n[tid]);
                  }
                  lb_n[tid][a++] = v;
                }
              } else {
                if (b == cap_f[tid]) {
                  cap_f[tid] = cap_f[tid] ? cap_f[tid] * 2 : 1024;
                  lb_f[tid] = (uint32_t *)realloc(lb_f[tid], sizeof(uint32_t) * cap_f[tid]);
                }
                lb_f[tid][b++] = v;
              }
            }
          }
        }
        tn[tid] = a;
        tf[tid] = b;
#pragma omp barrier
#pragma omp single
        {
          uint64_t s = 0, s2 = nfar;
          for (int t = 0; t < T; t++) {
            uint64_t c = tn[t], c2 = tf[t];
            tn[t] = s;
            tf[t] = s2;
            s += c;
            s2 += c2;
          }
          tn[T] = s;
          tf[T] = s2;
          if (s > ncap) {
            while (ncap < s) ncap *= 2;
            free(near);
            near = (uint32_t *)malloc(sizeof(uint32_t) * ncap);
          }
          if (s2 > fcap) {
            while (fcap < s2) fcap *= 2;
            far = (uint32_t *)realloc(far, sizeof(uint32_t) * fcap);
          }
        }
        memcpy(near + tn[tid], lb_n[tid], a * sizeof(uint32_t));
        memcpy(far + tf[tid], lb_f[tid], b * sizeof(uint32_t));
      }
      nn = tn[T];
      nfar = tf[T];
    }
    if (!nfar) break;
    /* advance the threshold to the smallest live far distance */
    uint32_t mn = OR_INFINIT;
    for (uint64_t i = 0; i < nfar; i++) {
      uint32_t dv = dist[far[i]];
      if (dv >= thr && dv < mn) mn = dv;
    }
    if (mn == OR_INFINIT) break;
    thr = ((uint64_t)mn / delta + 1) * delta;
    round++;
    uint64_t k = 0, m = 0;
    for (uint64_t i = 0; i < nfar; i++) {
      uint32_t v = far[i];
      uint32_t dv = dist[v];
      if (dv < thr - delta) continue; /* settled in an earlier bucket: stale */
      if (dv < thr) {
        if (stamp[v] != round) {
          stamp[v] = round;
          if (k == ncap) {
            ncap *= 2;
            near = (uint32_t *)realloc(near, sizeof(uint32_t) * ncap);
          }
          near[k++] = v;
        }
      } else {
        far[m++] = v;
      }
    }
    nn = k;
    nfar = m;
  }
  double t1 = now_ms();
  for (int t = 0; t < T; t++) {
    free(lb_n[t]);
    free(lb_f[t]);
  }
  free(lb_n);
  free(lb_f);
  free(cap_n);
  free(cap_f);
  free(tn);
  free(tf);
  free(near);
  free(far);
  free(stamp);
  return t1 - t0;
}

/* ------------------------------------------------------------------ */
double bl_pagerank(uint64_t V, const uint64_t *off, const uint32_t *col, float alpha,
                   int iters, float *rank, int nthreads) {
  int T = nth(nthreads);
  float *c = (float *)malloc(sizeof(float) * (V + 1));
  float base = (float)((1.0 - (double)alpha) / (double)V);
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t v = 0; v < V; v++) rank[v] = base;
  double t0 = now_ms();
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for num_threads(T) schedule(static)
    for (uint64_t u = 0; u < V; u++) {
      uint64_t d = off[u + 1] - off[u];
      c[u] = d ? rank[u] / (float)d : 0.0f;
    }
#pragma omp parallel for num_threads(T) schedule(dynamic, 1024)
    for (uint64_t v = 0; v < V; v++) {
      float s = 0.0f;
      for (uint64_t e = off[v]; e < off[v + 1]; e++) s += c[col[e]];
      rank[v] = base + alpha * s;
    }
  }
  double t1 = now_ms();
  free(c);
  return t1 - t0;
}

/* ------------------------------------------------------------------ */
static inline uint32_t ld(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

double bl_cc(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t *label,
             int nthreads) {
  int T = nth(nthreads);
  double t0 = now_ms();
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t v = 0; v < V; v++) label[v] = (uint32_t)v;
#pragma omp parallel for num_threads(T) schedule(dynamic, 256)
  for (uint64_t u = 0; u < V; u++) {
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t v = col[e];
      if (v >= u) continue;
      uint32_t p1 = ld(&label[u]), p2 = ld(&label[v]);
      while (p1 != p2) {
        uint32_t hi = p1 > p2 ? p1 : p2, lo = p1 > p2 ? p2 : p1;
        uint32_t ph = ld(&label[hi]);
        if (ph == lo) break;
        uint32_t exp = hi;
        if (ph == hi &&
            __atomic_compare_exchange_n(&label[hi], &exp, lo, 0, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED))
          break;
        p1 = ld(&label[ld(&label[hi])]);
        p2 = ld(&label[lo]);
      }
    }
  }
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t v = 0; v < V; v++) {
    uint32_t x = ld(&label[v]);
    while (x != ld(&label[x])) x = ld(&label[x]);
    label[v] = x;
  }
  double t1 = now_ms();
  return t1 - t0;
}
