/* Test infrastructure: CPU oracle for the parallel Zhang-Shasha tree edit
 * distance. Only tests, the smoke check and the bench's CPU-baseline leg use
 * this; the product path never links or calls it.
 *
 * Trees are preorder parent arrays (parent[0] = -1, parent[i] < i) with int32
 * label ids; equal ids are equal labels. TD outputs are m x n row-major in
 * preorder: td[i*n + j] = TED(subtree_x(i), subtree_y(j)) under unit costs.
 *
 * Parity status of each function (what pins it) is listed in the header of
 * xted_oracle.cpp and zs_textbook.c. */
#ifndef XTED_ORACLE_H
#define XTED_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Restatement of X-TED_CPU (xted_oracle.cpp). */
int oracle_orl(const int32_t* parent, int32_t n, int32_t* orl);
int oracle_keyroots(const int32_t* orl, int32_t n, int32_t* kr, int32_t* K);
int oracle_keyroot_depth(const int32_t* orl, const int32_t* kr, int32_t K, int32_t* depth);
/* Full level-synchronous computation. Returns TD[0][0] (the distance), or the
 * empty-tree answer (m==0 -> n, n==0 -> m), or -1 on invalid input. td may be
 * NULL. compute_ms (may be NULL) receives the level-loop time, the interval
 * the reference's own timer covers. */
int oracle_ted(const int32_t* x_parent, const int32_t* x_label, int32_t m,
               const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t threads,
               int32_t* td, double* compute_ms);
/* Non-unit costs (no counterpart in the reference, which hard-codes unit
 * costs, X-TED_CPU/TED.cpp:111,115,122): labels in [0, alphabet), del[a] and
 * ins[a] for deleting / inserting a node labelled a, ren[a*alphabet + b] for
 * relabelling a to b. A NULL costs pointer means unit costs. */
typedef struct {
  int32_t alphabet;
  const int32_t* del;
  const int32_t* ins;
  const int32_t* ren;
} oracle_costs;
/* oracle_ted under a cost model; the empty-tree answers become the sum of
 * the other tree's insertion (deletion) costs. */
int oracle_ted_costs(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                     const int32_t* y_parent, const int32_t* y_label, int32_t n,
                     const oracle_costs* costs, int32_t threads, int32_t* td,
                     double* compute_ms);
/* Level of every table (k, l): depth_x[k] + depth_y[l], K*L entries. */
int oracle_levels(const int32_t* x_parent, int32_t m, const int32_t* y_parent, int32_t n,
                  int32_t* levels, int32_t* K, int32_t* L);

/* Textbook postorder Zhang-Shasha (zs_textbook.c), independent of the
 * reference's preorder formulation. Same output convention. */
int zs_textbook(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t* td);
int zs_textbook_costs(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                      const int32_t* y_parent, const int32_t* y_label, int32_t n,
                      const oracle_costs* costs, int32_t* td);

#ifdef __cplusplus
}
#endif
#endif
