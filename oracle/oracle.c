/*
 * oracle.c -- serial CPU restatement of the Grus traversal semantics.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  See oracle.h for what pins each
 * function.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[256];
const char *or_last_error(void) { return g_err; }
static int fail(const char *m) {
  snprintf(g_err, sizeof g_err, "%s", m);
  return -1;
}

/* ------------------------------------------------------------------ */
/* .gr v1 reader -- include/graph_loader.cuh:65-159                    */
/* ------------------------------------------------------------------ */
int or_gr_header(const char *path, uint64_t *V, uint64_t *E, uint64_t *ety) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail("cannot open .gr file");
  uint64_t h[4];
  size_t n = fread(h, sizeof(uint64_t), 4, f); /* graph_loader.cuh:71-74 */
  fclose(f);
  if (n != 4) return fail("short .gr header");
  if (h[0] != 1) return fail("unknown .gr version"); /* :75-76 */
  *ety = h[1];
  *V = h[2];
  *E = h[3];
  return 0;
}

int or_gr_read(const char *path, uint64_t *off, uint32_t *col, uint32_t *w) {
  uint64_t V, E, ety;
  if (or_gr_header(path, &V, &E, &ety)) return -1;
  FILE *f = fopen(path, "rb");
  if (!f) return fail("cannot open .gr file");
  fseek(f, 32, SEEK_SET);
  off[0] = 0; /* the reference leaves xadj[0] to zeroed managed memory */
  if (V && fread(off + 1, sizeof(uint64_t), V, f) != V) { /* :114-120 */
    fclose(f);
    return fail("short .gr offsets");
  }
  if (E && fread(col, sizeof(uint32_t), E, f) != E) { /* :122-131 */
    fclose(f);
    return fail("short .gr destinations");
  }
  if (w) {
    if (ety) {
      if (E % 2) fseek(f, 4, SEEK_CUR); /* :139-141 pad */
      if (E && fread(w, sizeof(uint32_t), E, f) != E) { /* :142-150 */
        fclose(f);
        return fail("short .gr weights");
      }
    } else {
      for (uint64_t i = 0; i < E; i++) w[i] = 1; /* :95-101 */
    }
  }
  fclose(f);
  if (off[V] != E) return fail("offsets do not end at M");
  for (uint64_t i = 0; i < V; i++)
    if (off[i] > off[i + 1]) return fail("offsets not monotone");
  for (uint64_t i = 0; i < E; i++)
    if (col[i] >= V) return fail("destination out of range");
  return 0;
}

/* ------------------------------------------------------------------ */
/* generators: this project's RMAT and grid specification            */
/* ------------------------------------------------------------------ */
/* splitmix64 output function (Steele, Lea, Flood 2014; public domain) */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
#define GOLDEN 0x9E3779B97F4A7C15ull

static void thresholds(double a, double b, double c, uint32_t t[3]) {
  t[0] = (uint32_t)(a * 4294967296.0);
  t[1] = (uint32_t)((a + b) * 4294967296.0);
  t[2] = (uint32_t)((a + b + c) * 4294967296.0);
}

/* sample i: for level l (0 = most significant bit) a 32-bit uniform r is
 * taken from h_j = mix64(key + ((i<<5)|j) * GOLDEN), j = l/2, low half for
 * even l, high half for odd l.  Quadrant: r<t0 -> (0,0), r<t1 -> (0,1),
 * r<t2 -> (1,0), else (1,1); first bit to u, second to v. */
static inline void rmat_one(int scale, const uint32_t t[3], uint64_t key, uint64_t i,
                            uint32_t *uo, uint32_t *vo) {
  uint32_t u = 0, v = 0;
  uint64_t h = 0;
  for (int l = 0; l < scale; l++) {
    if ((l & 1) == 0) h = mix64(key + (((i << 5) | (uint64_t)(l >> 1)) * GOLDEN));
    uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
    uint32_t bu, bv;
    if (r < t[0]) { bu = 0; bv = 0; }
    else if (r < t[1]) { bu = 0; bv = 1; }
    else if (r < t[2]) { bu = 1; bv = 0; }
    else { bu = 1; bv = 1; }
    u = (u << 1) | bu;
    v = (v << 1) | bv;
  }
  *uo = u;
  *vo = v;
}

void or_rmat_samples(int scale, double a, double b, double c, uint64_t seed,
                     uint64_t i0, uint64_t i1, uint32_t *u, uint32_t *v) {
  uint32_t t[3];
  thresholds(a, b, c, t);
  uint64_t key = mix64(seed ^ 0x6A09E667F3BCC909ull);
  for (uint64_t i = i0; i < i1; i++) rmat_one(scale, t, key, i, &u[i - i0], &v[i - i0]);
}

uint32_t or_edge_weight(uint32_t u, uint32_t v, uint32_t wlo, uint32_t whi,
                        uint64_t wseed) {
  uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
  uint64_t key = mix64(wseed ^ 0xBB67AE8584CAA73Bull);
  uint64_t h = mix64(key + ((lo << 32) | hi));
  return wlo + (uint32_t)(h % (uint64_t)(whi - wlo + 1));
}

/* LSD radix sort of u64 keys, 16-bit digits; skips passes whose digit is
 * constant. */
static void radix_sort_u64(uint64_t *a, uint64_t n, uint64_t *tmp) {
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * 65536);
  for (int pass = 0; pass < 4; pass++) {
    int sh = pass * 16;
    memset(cnt, 0, sizeof(uint64_t) * 65536);
    for (uint64_t i = 0; i < n; i++) cnt[(a[i] >> sh) & 0xFFFF]++;
    int trivial = 0;
    for (int d = 0; d < 65536; d++)
      if (cnt[d] == n) trivial = 1;
    if (trivial) continue;
    uint64_t s = 0;
    for (int d = 0; d < 65536; d++) {
      uint64_t c = cnt[d];
      cnt[d] = s;
      s += c;
    }
    for (uint64_t i = 0; i < n; i++) tmp[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
    memcpy(a, tmp, n * sizeof(uint64_t));
  }
  free(cnt);
}

/* builds the symmetrized, deduplicated, self-loop-free sorted pair list */
static uint64_t *rmat_pairs(int scale, int ef, double a, double b, double c,
                            uint64_t seed, uint64_t *n_out) {
  uint64_t M = (uint64_t)ef << scale;
  uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * 2 * M + 8);
  uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * 2 * M + 8);
  if (!k || !tmp) {
    free(k);
    free(tmp);
    return NULL;
  }
  uint32_t t[3];
  thresholds(a, b, c, t);
  uint64_t key = mix64(seed ^ 0x6A09E667F3BCC909ull);
  uint64_t n = 0;
  for (uint64_t i = 0; i < M; i++) {
    uint32_t u, v;
    rmat_one(scale, t, key, i, &u, &v);
    if (u == v) continue;
    k[n++] = ((uint64_t)u << 32) | v;
    k[n++] = ((uint64_t)v << 32) | u;
  }
  radix_sort_u64(k, n, tmp);
  free(tmp);
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; i++)
    if (i == 0 || k[i] != k[i - 1]) k[m++] = k[i];
  *n_out = m;
  return k;
}

int64_t or_gen_rmat_count(int scale, int ef, double a, double b, double c,
                          uint64_t seed) {
  uint64_t m;
  uint64_t *k = rmat_pairs(scale, ef, a, b, c, seed, &m);
  if (!k) return fail("out of memory");
  free(k);
  return (int64_t)m;
}

int or_gen_rmat_fill(int scale, int ef, double a, double b, double c, uint64_t seed,
                     uint32_t wlo, uint32_t whi, uint64_t wseed, uint64_t *off,
                     uint32_t *col, uint32_t *w) {
  uint64_t m;
  uint64_t *k = rmat_pairs(scale, ef, a, b, c, seed, &m);
  if (!k) return fail("out of memory");
  uint64_t V = 1ull << scale;
  uint64_t e = 0;
  for (uint64_t r = 0; r <= V; r++) {
    while (e < m && (k[e] >> 32) < r) e++;
    off[r] = e;
  }
  for (uint64_t i = 0; i < m; i++) {
    col[i] = (uint32_t)k[i];
    if (w) w[i] = or_edge_weight((uint32_t)(k[i] >> 32), (uint32_t)k[i], wlo, whi, wseed);
  }
  free(k);
  return 0;
}

uint64_t or_gen_grid_count(uint32_t R, uint32_t C) {
  return 2ull * ((uint64_t)R * (C - 1) + (uint64_t)C * (R - 1));
}

void or_gen_grid_fill(uint32_t R, uint32_t C, uint32_t wlo, uint32_t whi, uint64_t wseed,
                      uint64_t *off, uint32_t *col, uint32_t *w) {
  uint64_t e = 0;
  for (uint64_t r = 0; r < R; r++)
    for (uint64_t c = 0; c < C; c++) {
      uint64_t id = r * C + c;
      off[id] = e;
      uint64_t nb[4];
      int n = 0;
      if (r > 0) nb[n++] = id - C;
      if (c > 0) nb[n++] = id - 1;
      if (c + 1 < C) nb[n++] = id + 1;
      if (r + 1 < R) nb[n++] = id + C;
      for (int j = 0; j < n; j++) {
        col[e] = (uint32_t)nb[j];
        if (w) w[e] = or_edge_weight((uint32_t)id, (uint32_t)nb[j], wlo, whi, wseed);
        e++;
      }
    }
  off[(uint64_t)R * C] = e;
}

uint64_t or_fnv64(const void *p, uint64_t n, uint64_t h) {
  const unsigned char *b = (const unsigned char *)p;
  if (h == 0) h = 0xcbf29ce484222325ull;
  for (uint64_t i = 0; i < n; i++) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

/* ------------------------------------------------------------------ */
/* BFS -- samples/bfs/bfs.cu:18-23, 51-60                              */
/* ------------------------------------------------------------------ */
void or_bfs(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
            uint32_t *level) {
  for (uint64_t i = 0; i < V; i++) level[i] = OR_INFINIT; /* BFSInit :21 */
  if (src >= V) return;
  uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
  uint64_t h = 0, t = 0;
  level[src] = 0;
  q[t++] = src;
  while (h < t) {
    uint32_t u = q[h++];
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t v = col[e];
      if (level[v] > level[u] + 1) { /* updater :54 */
        level[v] = level[u] + 1;
        q[t++] = v;
      }
    }
  }
  free(q);
}

/* BFS certificate on a symmetric graph: the levels are the hop distances
 * from src if and only if (1) level[src] = 0 and no other vertex has level 0,
 * (2) every neighbour of a reached vertex u is reached with level <= level[u]
 * + 1 (so level <= distance, along a shortest path), and (3) every reached
 * v != src has a neighbour at level[v] - 1 (so a path of level[v] edges
 * leads back to src: level >= distance).  O(E), parallel over vertices, for
 * graphs whose serial queue BFS is slow.  Returns 0 when valid, else the
 * number of the failed condition, with the vertex in *bad. */
int or_bfs_certify(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
                   const uint32_t *level, uint64_t *bad) {
  int code = 0;
  uint64_t where = 0;
  if (src >= V || level[src] != 0) {
    *bad = src;
    return 1;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V; v++) {
    const uint32_t lv = level[v];
    int c = 0;
    if (lv == OR_INFINIT) continue; /* unreached: condition 2 covers its neighbours */
    if (lv >= V || (lv == 0 && v != src)) {
      c = 1;
    } else {
      int parent = (v == src);
      for (uint64_t e = off[v]; e < off[v + 1]; e++) {
        const uint32_t lu = level[col[e]];
        if (lu == OR_INFINIT || lu > lv + 1) {
          c = 2;
          break;
        }
        if (lu + 1 == lv) parent = 1;
      }
      if (!c && !parent) c = 3;
    }
    if (c) {
#pragma omp critical
      {
        if (!code || v < where) {
          code = c;
          where = v;
        }
      }
    }
  }
  *bad = where;
  return code;
}

/* ------------------------------------------------------------------ */
/* SSSP -- samples/sssp/sssp.cu:18-23, 51-60; Dijkstra with binary heap */
/* ------------------------------------------------------------------ */
typedef struct {
  uint64_t d;
  uint32_t v;
} hitem;
static void heap_push(hitem *hp, uint64_t *n, hitem x) {
  uint64_t i = (*n)++;
  while (i) {
    uint64_t p = (i - 1) / 2;
    if (hp[p].d <= x.d) break;
    hp[i] = hp[p];
    i = p;
  }
  hp[i] = x;
}
static hitem heap_pop(hitem *hp, uint64_t *n) {
  hitem top = hp[0], x = hp[--(*n)];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    uint64_t md = x.d;
    if (l < *n && hp[l].d < md) { m = l; md = hp[l].d; }
    if (r < *n && hp[r].d < md) { m = r; }
    if (m == i) break;
    hp[i] = hp[m];
    i = m;
  }
  hp[i] = x;
  return top;
}

int or_sssp(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w,
            uint32_t src, uint32_t *dist) {
  for (uint64_t i = 0; i < V; i++) dist[i] = OR_INFINIT; /* SSSPInit :21 */
  if (src >= V) return 0;
  uint64_t E = off[V];
  hitem *hp = (hitem *)malloc(sizeof(hitem) * (E + 2));
  uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * V);
  for (uint64_t i = 0; i < V; i++) d[i] = UINT64_MAX;
  uint64_t n = 0;
  int rc = 0;
  d[src] = 0;
  heap_push(hp, &n, (hitem){0, src});
  while (n) {
    hitem x = heap_pop(hp, &n);
    if (x.d != d[x.v]) continue;
    for (uint64_t e = off[x.v]; e < off[x.v + 1]; e++) {
      uint64_t nd = x.d + (w ? w[e] : 1u);
      uint32_t v = col[e];
      if (nd < d[v]) {
        d[v] = nd;
        heap_push(hp, &n, (hitem){nd, v});
      }
    }
  }
  for (uint64_t i = 0; i < V; i++) {
    if (d[i] == UINT64_MAX) continue;
    if (d[i] >= OR_INFINIT) rc = fail("distance overflows u32");
    dist[i] = (uint32_t)d[i];
  }
  free(hp);
  free(d);
  return rc;
}

/* ------------------------------------------------------------------ */
/* CC -- samples/cc/cc.cu:35-40, 69-79; union-find, label = min id      */
/* ------------------------------------------------------------------ */
static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}
void or_cc(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t *label) {
  uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
  for (uint64_t i = 0; i < V; i++) p[i] = (uint32_t)i;
  for (uint64_t u = 0; u < V; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t a = uf_find(p, (uint32_t)u), b = uf_find(p, col[e]);
      if (a == b) continue;
      if (a < b) p[b] = a; /* root is always the smaller id */
      else p[a] = b;
    }
  for (uint64_t i = 0; i < V; i++) label[i] = uf_find(p, (uint32_t)i);
  free(p);
}

/* ------------------------------------------------------------------ */
/* PageRank FP64 -- samples/pagerank/pr.cu:17-43, 78-90 (Jacobi form)   */
/* ------------------------------------------------------------------ */
void or_pagerank(uint64_t V, const uint64_t *in_off, const uint32_t *in_col,
                 const uint64_t *out_off, double alpha, int iters, double *rank) {
  if (V == 0) return;
  double base = (1.0 - alpha) / (double)V;
  double *c = (double *)malloc(sizeof(double) * V);
  for (uint64_t v = 0; v < V; v++) rank[v] = base; /* pr_init :26 (scaled) */
  /* The loops run in parallel over vertices; each rank's sum keeps the
   * serial in-edge order, so the result does not depend on the thread count
   * (the full-size tests run it on 530 M edges for 20 iterations). */
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(static)
    for (uint64_t u = 0; u < V; u++) {
      uint64_t d = out_off[u + 1] - out_off[u];
      c[u] = d ? rank[u] / (double)d : 0.0; /* pr.cu:83 divide by out-degree */
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t v = 0; v < V; v++) {
      double s = 0.0;
      for (uint64_t e = in_off[v]; e < in_off[v + 1]; e++) s += c[in_col[e]];
      rank[v] = base + alpha * s;
    }
  }
  free(c);
}

/* ------------------------------------------------------------------ */
/* transpose with weights -- include/graph.cuh:273-303                 */
/* ------------------------------------------------------------------ */
void or_transpose(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w,
                  uint64_t *t_off, uint32_t *t_col, uint32_t *t_w) {
  uint64_t E = off[V];
  for (uint64_t i = 0; i <= V; i++) t_off[i] = 0;
  for (uint64_t e = 0; e < E; e++) t_off[col[e] + 1]++;
  for (uint64_t i = 0; i < V; i++) t_off[i + 1] += t_off[i];
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * (V ? V : 1));
  for (uint64_t i = 0; i < V; i++) pos[i] = t_off[i];
  for (uint64_t u = 0; u < V; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint64_t d = pos[col[e]]++;
      t_col[d] = (uint32_t)u;
      if (t_w && w) t_w[d] = w[e];
    }
  free(pos);
}
