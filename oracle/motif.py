"""Oracle restatement of the reference motif reader (test infrastructure only).

Follows:
  * ``TemporalGraph::ReadGraph`` (src/core/cpu/Graph.cc:198-259): each line is
    read as ``u v t`` by stream extraction, the edges are stable-sorted by
    ``t`` and inserted in that order; ``insert`` (Graph.cc:72-96) renames nodes
    by first appearance, source before destination.
  * ``HostMotifData::HostMotifData`` (src/core/cpu/data.cc:411-449): the
    per-edge pre-decoding into ``MotifEdgeInfoV1`` records for edges 1..M-1.

Pinned by: the compiled reference reader in ``oracle/_ref`` (built from the
reference Graph.cc by ``oracle/Makefile``) on every reference motif fixture and
on random motif files (tests/test_oracle.py::test_motif_reader_matches_reference_build).
"""
from dataclasses import dataclass
from typing import List, Tuple


@dataclass(frozen=True)
class MInfo:
    """One ``MotifEdgeInfoV1`` record (include/data.h:33-40), minus pointers."""
    baseNode: int
    constraintNode: int
    mappedNodes: int
    io: int
    newNode: bool


def read_motif_text(text: str) -> List[Tuple[int, int, int]]:
    """Parse ``u v t`` lines (blank lines skipped), stable sort by t,
    rename nodes by first appearance (u before v).  Returns renamed edges."""
    raw = []
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        raw.append((int(f[0]), int(f[1]), int(f[2])))
    # Graph.cc:234 std::stable_sort on t
    raw = sorted(raw, key=lambda e: e[2])
    ids = {}
    out = []
    for (u, v, t) in raw:
        for n in (u, v):
            if n not in ids:
                ids[n] = len(ids)
        out.append((ids[u], ids[v], t))
    return out


def read_motif(path: str) -> List[Tuple[int, int, int]]:
    with open(path, "r") as f:
        return read_motif_text(f.read())


def predecode(edges: List[Tuple[int, int, int]]) -> List[MInfo]:
    """``HostMotifData`` pre-decoding (src/core/cpu/data.cc:422-447).

    Raises ValueError for an edge with two unmapped endpoints, which the
    reference leaves undefined (its ``baseNode`` would be unmapped)."""
    node_max = 1
    out = []
    for i in range(1, len(edges)):
        u, v, _ = edges[i]
        if u > node_max and v > node_max:
            raise ValueError("motif edge %d has two unmapped endpoints" % i)
        if u > node_max:
            out.append(MInfo(v, -u, node_max + 1, 0, True))
            node_max += 1
        elif v > node_max:
            out.append(MInfo(u, -v, node_max + 1, 1, True))
            node_max += 1
        else:
            out.append(MInfo(v, u, node_max + 1, -1, False))
    return out


def num_nodes(edges) -> int:
    return 1 + max(max(u, v) for u, v, _ in edges)
