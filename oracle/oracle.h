/*
 * oracle.h -- CPU restatement of the Grus traversal engine's semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (grus_framework_amd/)
 * may link, load or call this code; only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only
 * as the checker / the timed CPU baseline, never as the thing measured.
 *
 * Every function names the reference code whose semantics it restates
 * (paths relative to the reference repository root).
 *
 * Parity pinning (what checks each function, beyond "same formula"):
 *   or_read_gr        hand-made byte fixtures (odd/even M, weighted and
 *                     unweighted) in tests/golden, written byte by byte in
 *                     the test, not by the product's writer.
 *   or_gen_rmat/grid  generator spec is this project's own (the reference
 *                     has no generator): pinned by committed CSR hashes,
 *                     quadrant statistics vs a/b/c/d, symmetry, sortedness,
 *                     and by the independent GPU generator.
 *   or_bfs            closed forms (path, star, cycle, complete, grid
 *                     corner = r+c) and brute-force all-pairs on tiny graphs.
 *   or_bfs_certify    accepts or_bfs's levels and scipy's hop distances on
 *                     RMAT and grids; rejects planted mistakes (a level one
 *                     too high or too low, a second level-0 vertex, a wrong
 *                     source, an unreached vertex beside a reached one).
 *   or_sssp           closed forms (unit-weight grid = Manhattan distance),
 *                     Floyd-Warshall on <= 64 vertices, O(E) certificate.
 *   or_cc             invariants (label constant on edges, label<=v,
 *                     label[label[v]]==label[v]) and brute-force closure.
 *   or_pagerank       closed forms: k-regular graphs stay uniform, isolated
 *                     vertex = (1-a)/V, 2-vertex path/star analytic iterates.
 */
#ifndef GRUS_ORACLE_H
#define GRUS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INFINIT 0xFFFFFFFEu /* include/common.cuh:52  UINT_MAX - 1 */

/* Errors: 0 ok, <0 failure (message in or_last_error()). */
const char *or_last_error(void);

/* ---- .gr v1 reader (include/graph_loader.cuh:65-159) ----
 * header u64[4] {version=1, sizeEdgeTy, N, M}; u64[N] end offsets;
 * u32[M] destinations; 4-byte pad when sizeEdgeTy!=0 and M is odd;
 * u32[M] weights when sizeEdgeTy!=0.  Offsets are kept as u64 and off[0]=0
 * (the reference truncates to u32 and never writes xadj[0]).  *w_out is set
 * to NULL when the file carries no weights. */
int or_gr_header(const char *path, uint64_t *V, uint64_t *E, uint64_t *size_edge_ty);
int or_gr_read(const char *path, uint64_t *off /*V+1*/, uint32_t *col /*E*/,
               uint32_t *w /*E or NULL*/);

/* ---- synthetic graphs (project spec; reference has none) ----
 * Two-phase API: *_count returns E, *_fill writes the CSR. */
int64_t or_gen_rmat_count(int scale, int edge_factor, double a, double b, double c,
                          uint64_t seed);
int or_gen_rmat_fill(int scale, int edge_factor, double a, double b, double c,
                     uint64_t seed, uint32_t wlo, uint32_t whi, uint64_t wseed,
                     uint64_t *off, uint32_t *col, uint32_t *w /*or NULL*/);
/* raw samples (u,v) for sample indices [i0,i1): used by statistics tests */
void or_rmat_samples(int scale, double a, double b, double c, uint64_t seed,
                     uint64_t i0, uint64_t i1, uint32_t *u, uint32_t *v);
uint32_t or_edge_weight(uint32_t u, uint32_t v, uint32_t wlo, uint32_t whi,
                        uint64_t wseed);
uint64_t or_gen_grid_count(uint32_t rows, uint32_t cols);
void or_gen_grid_fill(uint32_t rows, uint32_t cols, uint32_t wlo, uint32_t whi,
                      uint64_t wseed, uint64_t *off, uint32_t *col, uint32_t *w);

/* FNV-1a 64 over the CSR arrays (little-endian bytes of off, col, w). */
uint64_t or_fnv64(const void *p, uint64_t nbytes, uint64_t h);

/* ---- apps ---- */
/* BFS levels: samples/bfs/bfs.cu:18-23 (init), :51-60 (updater: level =
 * iteration+1 on first discovery).  Unreached = OR_INFINIT. */
int or_bfs_certify(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
                   const uint32_t *level, uint64_t *bad);
void or_bfs(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
            uint32_t *level);
/* SSSP distances: samples/sssp/sssp.cu:18-23, :51-60 (relax d[v] >
 * d[u]+w[e]); fixed point computed by binary-heap Dijkstra.  w==NULL means
 * unit weights (graph_loader.cuh:95-101).  Returns 0, or -1 if a distance
 * would reach OR_INFINIT. */
int or_sssp(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w,
            uint32_t src, uint32_t *dist);
/* CC labels: samples/cc/cc.cu:35-40 (label=id), :69-79 (atomicMin
 * propagation) -> fixed point = minimum vertex id of each component of the
 * symmetric graph.  Computed by union-find. */
void or_cc(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t *label);
/* PageRank, FP64, Jacobi form of the reference's delta recurrence
 * (samples/pagerank/pr.cu:17-43, :78-90; ALPHA include/common.cuh:32),
 * normalised by V:  r0 = (1-a)/V;  r_{k+1}(v) = (1-a)/V + a * sum over
 * in-edges (u,v) of r_k(u)/outdeg(u).  Dangling mass is dropped, as in the
 * reference.  in_off/in_col is the CSC (== CSR for symmetric graphs);
 * outdeg from the CSR. */
void or_pagerank(uint64_t V, const uint64_t *in_off, const uint32_t *in_col,
                 const uint64_t *out_off, double alpha, int iters, double *rank);

/* CSR -> CSC with weights carried (include/graph.cuh:273-303; the
 * reference drops weights at :294). */
void or_transpose(uint64_t V, const uint64_t *off, const uint32_t *col,
                  const uint32_t *w, uint64_t *t_off, uint32_t *t_col, uint32_t *t_w);

/* ---- OpenMP CPU baselines (same semantics; timed by bench.py) ----
 * Each returns elapsed milliseconds of the traversal proper and writes the
 * same output as the matching serial oracle.  nthreads<=0 = all cores. */
double bl_bfs(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t src,
              uint32_t *level, int nthreads);
double bl_sssp(uint64_t V, const uint64_t *off, const uint32_t *col, const uint32_t *w,
               uint32_t src, uint32_t delta, uint32_t *dist, int nthreads);
double bl_pagerank(uint64_t V, const uint64_t *off, const uint32_t *col, float alpha,
                   int iters, float *rank, int nthreads);
double bl_cc(uint64_t V, const uint64_t *off, const uint32_t *col, uint32_t *label,
             int nthreads);
int bl_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
