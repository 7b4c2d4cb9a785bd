/* Test infrastructure (oracle). Textbook Zhang-Shasha (K. Zhang and
 * D. Shasha, "Simple fast algorithms for the editing distance between trees
 * and related problems", SIAM J. Comput. 18(6), 1989): postorder numbering,
 * leftmost-leaf descendants l(i), LR-keyroots in increasing postorder, and the
 * forest-distance table fd per keyroot pair. It shares no code and no
 * numbering with the reference's preorder / outermost-right-leaf formulation,
 * so it serves as an independent check of the oracle restatement.
 *
 * Parity pins: brute force over all edit mappings on trees of <= 6 nodes and
 * the Levenshtein closed forms (tests/test_oracle.py). */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int min3i(int a, int b, int c) {
  int m = a < b ? a : b;
  return m < c ? m : c;
}

/* Postorder view of a preorder tree: post[pre] and, indexed by postorder
 * position p (1-based), the preorder node, leftmost-leaf position and label. */
typedef struct {
  int n;
  int* pre_of;  /* [1..n] -> preorder id */
  int* lml;     /* [1..n] -> postorder position of the leftmost leaf */
  int* lab;     /* [1..n] */
  int* kr;      /* LR keyroots, increasing postorder */
  int nkr;
} Post;

static int build_post(const int32_t* parent, const int32_t* label, int n, Post* P) {
  int *first_child = malloc(sizeof(int) * n), *next_sib = malloc(sizeof(int) * n);
  int *last_child = malloc(sizeof(int) * n), *post = malloc(sizeof(int) * n);
  int *stack = malloc(sizeof(int) * (n + 1)), *it = malloc(sizeof(int) * n);
  P->n = n;
  P->pre_of = malloc(sizeof(int) * (n + 1));
  P->lml = malloc(sizeof(int) * (n + 1));
  P->lab = malloc(sizeof(int) * (n + 1));
  P->kr = malloc(sizeof(int) * (n + 1));
  for (int i = 0; i < n; ++i) first_child[i] = next_sib[i] = last_child[i] = -1;
  for (int i = 1; i < n; ++i) {
    int p = parent[i];
    if (first_child[p] < 0) first_child[p] = i;
    else next_sib[last_child[p]] = i;
    last_child[p] = i;
  }
  /* iterative postorder */
  int sp = 0, pos = 0;
  stack[sp] = 0;
  it[0] = first_child[0];
  while (sp >= 0) {
    int v = stack[sp];
    int c = it[v];
    if (c >= 0) {
      it[v] = next_sib[c];
      stack[++sp] = c;
      it[c] = first_child[c];
    } else {
      post[v] = ++pos;
      P->pre_of[pos] = v;
      --sp;
    }
  }
  for (int p = 1; p <= n; ++p) {
    int v = P->pre_of[p];
    while (first_child[v] >= 0) v = first_child[v];
    P->lml[p] = post[v];
    P->lab[p] = label[P->pre_of[p]];
  }
  /* keyroot: highest postorder node for each distinct leftmost leaf */
  int* seen = calloc(n + 2, sizeof(int));
  P->nkr = 0;
  for (int p = n; p >= 1; --p) {
    if (!seen[P->lml[p]]) {
      seen[P->lml[p]] = 1;
      P->kr[P->nkr++] = p;
    }
  }
  /* reverse to increasing order */
  for (int a = 0, b = P->nkr - 1; a < b; ++a, --b) {
    int t = P->kr[a];
    P->kr[a] = P->kr[b];
    P->kr[b] = t;
  }
  free(seen);
  free(first_child);
  free(next_sib);
  free(last_child);
  free(post);
  free(stack);
  free(it);
  return 0;
}

static void free_post(Post* P) {
  free(P->pre_of);
  free(P->lml);
  free(P->lab);
  free(P->kr);
}

int zs_textbook(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t* td) {
  return zs_textbook_costs(x_parent, x_label, m, y_parent, y_label, n, NULL, td);
}

/* Cost of deleting / inserting the node at postorder position p, and of
 * relabelling a onto b. */
static int c_del(const oracle_costs* C, const Post* A, int p) { return C ? C->del[A->lab[p]] : 1; }
static int c_ins(const oracle_costs* C, const Post* B, int p) { return C ? C->ins[B->lab[p]] : 1; }
static int c_ren(const oracle_costs* C, const Post* A, int i, const Post* B, int j) {
  return C ? C->ren[(size_t)A->lab[i] * C->alphabet + B->lab[j]] : (A->lab[i] != B->lab[j]);
}

int zs_textbook_costs(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                      const int32_t* y_parent, const int32_t* y_label, int32_t n,
                      const oracle_costs* C, int32_t* td) {
  if (m == 0 || n == 0) {
    int s = 0;
    for (int j = 0; j < n; ++j) s += C ? C->ins[y_label[j]] : 1;
    for (int i = 0; i < m; ++i) s += C ? C->del[x_label[i]] : 1;
    return s;
  }
  Post A, B;
  build_post(x_parent, x_label, m, &A);
  build_post(y_parent, y_label, n, &B);
  /* treedist indexed by postorder positions 1..m x 1..n */
  int* T = malloc(sizeof(int) * (size_t)(m + 1) * (n + 1));
  int* fd = malloc(sizeof(int) * (size_t)(m + 1) * (n + 1));
  const int W = n + 1;
  for (int a = 0; a < A.nkr; ++a) {
    for (int b = 0; b < B.nkr; ++b) {
      const int i = A.kr[a], j = B.kr[b];
      const int li = A.lml[i], lj = B.lml[j];
      /* fd rows li-1..i, cols lj-1..j (absolute postorder positions) */
      fd[(li - 1) * W + (lj - 1)] = 0;
      for (int i1 = li; i1 <= i; ++i1)
        fd[i1 * W + (lj - 1)] = fd[(i1 - 1) * W + (lj - 1)] + c_del(C, &A, i1);
      for (int j1 = lj; j1 <= j; ++j1)
        fd[(li - 1) * W + j1] = fd[(li - 1) * W + j1 - 1] + c_ins(C, &B, j1);
      for (int i1 = li; i1 <= i; ++i1) {
        for (int j1 = lj; j1 <= j; ++j1) {
          const int del = fd[(i1 - 1) * W + j1] + c_del(C, &A, i1);
          const int ins = fd[i1 * W + j1 - 1] + c_ins(C, &B, j1);
          if (A.lml[i1] == li && B.lml[j1] == lj) {
            const int ren = fd[(i1 - 1) * W + j1 - 1] + c_ren(C, &A, i1, &B, j1);
            const int v = min3i(del, ins, ren);
            fd[i1 * W + j1] = v;
            T[i1 * W + j1] = v;
          } else {
            const int v = min3i(del, ins,
                                fd[(A.lml[i1] - 1) * W + (B.lml[j1] - 1)] + T[i1 * W + j1]);
            fd[i1 * W + j1] = v;
          }
        }
      }
    }
  }
  const int dist = T[m * W + n];
  if (td) {
    for (int p = 1; p <= m; ++p)
      for (int q = 1; q <= n; ++q) td[(size_t)A.pre_of[p] * n + B.pre_of[q]] = T[p * W + q];
  }
  free(T);
  free(fd);
  free_post(&A);
  free_post(&B);
  return dist;
}
