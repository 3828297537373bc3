/*
 * tm_oracle.c -- CPU restatement of the reference temporal-motif DFS.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker and the CPU baseline
 * ("port"), never the product: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Follows, line by line in meaning:
 *   plugins/GPUWorkerU1B.cu:10-263       per-root DFS, register stack, count
 *   system/templates/common/resolveMi.snp:1-29   list choice (io, ties -> out)
 *   system/templates/funcs/reduceEarlyIdx.snp:1-20  first slot with id > i_g
 *   system/templates/common/nextMatchInTask.snp:1-21 scan, stop on t > tl
 *   system/templates/common/structuralConstraint.snp:1-6 and
 *   system/tmine/KernelGen.py:426-458      same / not-same node checks
 *   system/templates/common/recovermi.snp:1-4    restore io on backtrack
 *   system/templates/dispatch/processFirstEdge.snp:1  roots with u==v skipped
 *   system/tmine/KernelGen.py:509-525      enumeration row = Euid[p_0..p_{M-1}]
 *
 * One deliberate difference: tl = t0 + delta is formed in 64-bit, where the
 * reference adds two ints (undefined on overflow; unreachable for int32 times
 * below 2^31 - delta).
 *
 * Work counters (per query, exact): R roots examined (end-beg), X1 / X2
 * resolveMi calls whose motif edge has io >= 0 / io == -1, P candidate probes
 * (iterations of the nextMatchInTask scan, terminating t > tl probe included).
 *
 * Pinned by: brute force from the match definition on random tiny graphs,
 * closed forms (parallel edges -> C(n,M), stars, cycles), and invariants
 * (root-range additivity, monotone in delta, relabel invariance) in
 * tests/test_oracle.py; the work counters by their restatement over the
 * enumerated prefix matches and the chosen lists' window counts in
 * tests/test_oracle_work.py.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int u, v, t; } oedge;
typedef struct { int baseNode, constraintNode, mappedNodes, io, newNode; } ominfo;

#define OMAXE 16
#define OMAXN 18

typedef struct {
  const oedge *Eg; int numeg;
  const int *inV, *inR, *outV, *outR;
  const ominfo *mi; int numem;
  long long delta;
  /* enumeration */
  const int *Euid; int *rows; long long cap; long long nrows;
} octx;

static void reduce_early_idx(const int *iS, int *beg, int end, int ig) {
  int b = *beg, e = end;
  if (e == b) { *beg = e; return; }
  if (iS[e - 1] <= ig) { *beg = e; return; }
  if (iS[b] > ig) return;
  while (b < e) {
    int p = b + (e - b) / 2;
    if (iS[p] > ig) e = p; else b = p + 1;
  }
  *beg = b;
}

/* resolveMi: pick the adjacency list for motif edge (level+1). */
static void resolve(const octx *c, int level, const int *m, int *beg, int *end,
                    const int **iS, int *io, int *cn, uint64_t *w) {
  const ominfo *mi = &c->mi[level];
  if (mi->io >= 0) {
    int base = m[mi->baseNode];
    const int *R = mi->io == 1 ? c->outR : c->inR;
    *beg = R[base]; *end = R[base + 1];
    *iS = mi->io == 1 ? c->outV : c->inV;
    *io = mi->io; *cn = mi->constraintNode;
    w[1]++;
  } else {
    int b0 = m[mi->baseNode], b1 = m[mi->constraintNode];
    int beg0 = c->inR[b0], end0 = c->inR[b0 + 1];
    int beg1 = c->outR[b1], end1 = c->outR[b1 + 1];
    if (end0 - beg0 < end1 - beg1) {
      *io = 0; *cn = mi->constraintNode; *beg = beg0; *end = end0; *iS = c->inV;
    } else {
      *io = 1; *cn = mi->baseNode; *beg = beg1; *end = end1; *iS = c->outV;
    }
    w[2]++;
  }
}

/* DFS from root tid; adds matches to *cnt and work to w[1..3]. */
static void dfs_root(octx *c, int tid, uint64_t *cnt, uint64_t *w, int enumerate) {
  const oedge *e0 = &c->Eg[tid];
  if (e0->u == e0->v) return;
  int m[OMAXN];
  int sb[OMAXE], se[OMAXE], sio[OMAXE], scn[OMAXE], eid[OMAXE];
  m[0] = e0->u; m[1] = e0->v;
  long long tl = (long long)e0->t + c->delta;
  int level = 0, beg, end, io, cn;
  const int *iS;
  eid[0] = tid;
  int ig = tid;
  resolve(c, 0, m, &beg, &end, &iS, &io, &cn, w);
  sio[0] = io; scn[0] = cn;
  reduce_early_idx(iS, &beg, end, ig);
  for (;;) {
    int found = 0, node = 0;
    const ominfo *mi = &c->mi[level];
    for (; beg < end; beg++) {
      w[3]++;
      int idx = iS[beg];
      const oedge *eg = &c->Eg[idx];
      if (eg->t > tl) { end = beg; break; }
      node = io ? eg->v : eg->u;
      int ok = 1;
      if (mi->constraintNode < 0) {
        for (int k = 0; k < mi->mappedNodes; k++) ok = ok && (m[k] != node);
      } else {
        ok = (m[cn] == node);
      }
      if (ok) { ig = idx; found = 1; beg++; break; }
    }
    if (found) {
      if (level == c->numem - 2) {
        (*cnt)++;
        if (enumerate) {
          long long slot = __atomic_fetch_add(&c->nrows, 1, __ATOMIC_RELAXED);
          if (slot < c->cap) {
            int *row = c->rows + slot * c->numem;
            for (int k = 0; k <= level; k++) row[k] = c->Euid[eid[k]];
            row[level + 1] = c->Euid[ig];
          }
        }
      } else {
        if (mi->newNode) m[mi->mappedNodes] = node;
        sb[level] = beg; se[level] = end;
        level++;
        eid[level] = ig;
        resolve(c, level, m, &beg, &end, &iS, &io, &cn, w);
        sio[level] = io; scn[level] = cn;
        reduce_early_idx(iS, &beg, end, ig);
      }
    } else {
      if (level) {
        level--;
        beg = sb[level]; end = se[level];
        io = sio[level]; cn = scn[level];
        iS = io ? c->outV : c->inV;
      } else {
        break;
      }
    }
  }
}

typedef struct {
  octx *c; int beg, end; int *next; int enumerate;
  uint64_t cnt; uint64_t w[4];
} oworker;

#define OCHUNK 4096

static void *worker_main(void *arg) {
  oworker *wk = (oworker *)arg;
  for (;;) {
    int s = __atomic_fetch_add(wk->next, OCHUNK, __ATOMIC_RELAXED);
    if (s >= wk->end) break;
    int e = s + OCHUNK < wk->end ? s + OCHUNK : wk->end;
    for (int r = s; r < e; r++) dfs_root(wk->c, r, &wk->cnt, wk->w, wk->enumerate);
  }
  return NULL;
}

static int run_all(octx *c, int beg, int end, int nthreads, int enumerate,
                   unsigned long long *count, unsigned long long *work) {
  if (c->numem < 2 || c->numem > OMAXE) return -1;
  for (int i = 0; i < c->numem - 1; i++)
    if (c->mi[i].mappedNodes + 1 > OMAXN) return -1;
  if (end < 0) end = c->numeg - c->numem + 1;
  if (beg < 0) beg = 0;
  if (end > c->numeg) end = c->numeg;
  if (nthreads < 1) nthreads = 1;
  int next = beg;
  oworker *wk = (oworker *)calloc((size_t)nthreads, sizeof(oworker));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int i = 0; i < nthreads; i++) {
    wk[i].c = c; wk[i].beg = beg; wk[i].end = end; wk[i].next = &next; wk[i].enumerate = enumerate;
  }
  if (nthreads == 1) {
    worker_main(&wk[0]);
  } else {
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, worker_main, &wk[i]);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  }
  unsigned long long cnt = 0, w[4] = {0, 0, 0, 0};
  for (int i = 0; i < nthreads; i++) {
    cnt += wk[i].cnt;
    for (int k = 1; k < 4; k++) w[k] += wk[i].w[k];
  }
  w[0] = end > beg ? (unsigned long long)(end - beg) : 0;
  *count = cnt;
  if (work) memcpy(work, w, sizeof(w));
  free(wk); free(th);
  return 0;
}

int tm_oracle_count(const int *Eg, int numeg, const int *inV, const int *inR,
                    const int *outV, const int *outR, const int *minfo5, int numem,
                    long long delta, int beg, int end, int nthreads,
                    unsigned long long *count, unsigned long long *work) {
  octx c;
  memset(&c, 0, sizeof(c));
  c.Eg = (const oedge *)Eg; c.numeg = numeg;
  c.inV = inV; c.inR = inR; c.outV = outV; c.outR = outR;
  c.mi = (const ominfo *)minfo5; c.numem = numem; c.delta = delta;
  return run_all(&c, beg, end, nthreads, 0, count, work);
}

/* Enumerates with one thread in root order; rows beyond cap are counted only. */
int tm_oracle_enum(const int *Eg, int numeg, const int *inV, const int *inR,
                   const int *outV, const int *outR, const int *minfo5, int numem,
                   long long delta, int beg, int end, const int *Euid,
                   long long cap, int *rows, unsigned long long *count) {
  octx c;
  memset(&c, 0, sizeof(c));
  c.Eg = (const oedge *)Eg; c.numeg = numeg;
  c.inV = inV; c.inR = inR; c.outV = outV; c.outR = outR;
  c.mi = (const ominfo *)minfo5; c.numem = numem; c.delta = delta;
  c.Euid = Euid; c.rows = rows; c.cap = cap; c.nrows = 0;
  return run_all(&c, beg, end, 1, 1, count, NULL);
}
