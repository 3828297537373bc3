"""Constraint-feature semantics on the CPU (test infrastructure only): a plain
depth-first search over the reference's candidate lists with label
predicates on motif nodes and edges, fine-grained temporal windows and a
filter over the matched labels, as the reference's generated worker applies
them (system/tmine/Query.py:136-308, system/tmine/KernelGen.py:467-477,
530-534; root checks in renderUniform*CheckFirst, per-candidate checks in
renderUniform*CheckLoop).  Pinned in tests/test_cons_check.py against the
oracle: without constraints, on edge lists restricted to the edges a uniform
predicate keeps, and by filtering the oracle's enumerated rows.

node_pred / edge_pred: {motif node or edge index: callable(label) -> bool};
temporal: gaps [g0, g1, ...] (edge j + 1 at most g_j after edge j, edge 1 at
most g0 after the root, every edge at most delta after the root from edge 2
on); filt: callable(edge_labels, node_labels) -> bool over a complete match;
EF: label per edge position, VF: label per (renamed) vertex; out_rows, when
given, collects the matches' edge positions (without anti edges).
"""
import numpy as np


def motif_structure(edges):
    """Per edge j >= 1: (kind, a, b, mapped) as the library's motif
    descriptor (0 new destination, 1 new source, 2 both mapped)."""
    out, n = [None], 2
    for j in range(1, len(edges)):
        a, b = edges[j]
        if a < n and b < n:
            out.append((2, a, b, n))
        elif a < n:
            out.append((0, a, b, n))
            n += 1
        else:
            out.append((1, a, b, n))
            n += 1
    return out, n


def anti_prev(anti, M):
    """Per edge j: the nearest earlier non-anti edge (Query.py:64-74)."""
    anti = list(anti) if anti else [False] * M
    prev = [0] * M
    for j in range(1, M):
        prev[j] = max(i for i in range(j) if not anti[i])
    return anti, prev


def count_constrained(g, edges, delta, EF, VF, node_pred=None, edge_pred=None, temporal=None, filt=None,
                      beg=0, end=None, anti=None, out_rows=None):
    node_pred, edge_pred = node_pred or {}, edge_pred or {}
    Eg, E = g["Eg"], g["E"]
    T = Eg[:, 2]
    outV, outR, inV, inR = g["outV"], g["outR"], g["inV"], g["inR"]
    M = len(edges)
    st, N = motif_structure(edges)
    end = E - M + 1 if end is None else end
    anti, prev = anti_prev(anti, M)
    assert not anti[0] and all(st[j][0] == 2 for j in range(1, M) if anti[j]), "anti edges join mapped nodes"
    npred = lambda i, x: node_pred[i](int(VF[x])) if i in node_pred else True
    epred = lambda j, p: edge_pred[j](int(EF[p])) if j in edge_pred else True
    total = 0

    def window(j, tj, t0):
        if temporal is None:
            return t0 + delta
        if j == 0:
            return tj + temporal[0]
        return min(tj + temporal[j], t0 + delta)

    def candidates(j, m, p):
        kind, a, b, _ = st[j]
        if kind == 0:
            x = m[a]
            for q in outV[outR[x]:outR[x + 1]]:
                if q > p:
                    yield int(q), int(Eg[q, 1])
        elif kind == 1:
            y = m[b]
            for q in inV[inR[y]:inR[y + 1]]:
                if q > p:
                    yield int(q), int(Eg[q, 0])
        else:
            x, y = m[a], m[b]
            for q in outV[outR[x]:outR[x + 1]]:
                if q > p and Eg[q, 1] == y:
                    yield int(q), -1

    def dfs(j, m, pos, t0):
        nonlocal total
        p = pos[prev[j]]
        tl = window(j - 1, int(T[p]), t0)
        kind, a, b, nm = st[j]
        if anti[j]:
            if any(T[q] <= tl and epred(j, q) for q, _ in candidates(j, m, p)):
                return
            if j == M - 1:
                if filt is None or filt([int(EF[x]) for x in pos + [p]], [int(VF[x]) for x in m]):
                    total += 1
            else:
                dfs(j + 1, m, pos + [p], t0)
            return
        for q, node in candidates(j, m, p):
            if T[q] > tl:
                break
            mm = m
            if kind != 2:
                if node in m or not npred(nm, node):
                    continue
                mm = m + [node]
            if not epred(j, q):
                continue
            if j == M - 1:
                if filt is None or filt([int(EF[x]) for x in pos + [q]], [int(VF[x]) for x in mm]):
                    total += 1
                    if out_rows is not None:
                        out_rows.append(tuple(pos + [q]))
            else:
                dfs(j + 1, mm, pos + [q], t0)

    for r in range(beg, end):
        u, v, t0 = (int(x) for x in Eg[r])
        if u == v or not epred(0, r) or not npred(0, u) or not npred(1, v):
            continue
        dfs(1, [u, v], [r], t0)
    return total


def match_rows(g, rows, edges):
    """The oracle's enumerated rows (Euid tuples) as (positions, node maps)."""
    inv = np.empty(g["E"], dtype=np.int64)
    inv[np.asarray(g["Euid"])] = np.arange(g["E"])
    st, N = motif_structure(edges)
    out = []
    for row in rows:
        pos = [int(inv[x]) for x in row]
        m = [int(g["Eg"][pos[0], 0]), int(g["Eg"][pos[0], 1])]
        for j in range(1, len(edges)):
            kind, a, b, nm = st[j]
            if kind == 0:
                m.append(int(g["Eg"][pos[j], 1]))
            elif kind == 1:
                m.append(int(g["Eg"][pos[j], 0]))
        out.append((pos, m))
    return out
