"""Oracle restatement of the reference edge-list loader and CSR build
(test infrastructure only).

Follows:
  * ``EdgeListLoader::createEdgeListCache`` (src/core/cpu/data.cc:175-253):
    fields parsed in file order, then ``ranges::stable_sort`` of (U,V,T,IDS)
    keyed on T (data.cc:246-248); IDS[p] is the original 0-based line number.
  * ``GraphDataLoader::createGraphData`` (src/core/gpu/data.cu:124-348):
    vertices renamed by first occurrence in the sorted list, a source before a
    destination on the same edge (data.cu:143-236); ``outEdgesV``/``inEdgesV``
    are sorted-edge positions grouped by renamed source/destination in
    ascending position order (stable sorts, data.cu:238-248); ``outEdgesR``/
    ``inEdgesR`` are the CSR offsets (data.cu:249-271); Eg = (u', v', t).

Pinned by: the closed-form checks in tests/test_oracle_prep.py (a stable sort
keeps file order on ties; renaming is a bijection in first-appearance order;
CSR lists are exactly the positions with that endpoint, ascending) and the
brute-force oracle, which works on the raw edge list and never calls this file.
"""
import numpy as np


def parse_edge_text(text: str):
    """Parse ``u v t`` lines (data.cc:94-111 reads three space-separated ints)."""
    arr = np.array(text.split(), dtype=np.int64)
    if arr.size % 3:
        raise ValueError("edge file does not hold whole u v t records")
    arr = arr.reshape(-1, 3).astype(np.int32)
    return arr[:, 0].copy(), arr[:, 1].copy(), arr[:, 2].copy()


def preprocess(U, V, T):
    """Return a dict with the reference host-graph arrays for the full range."""
    U = np.asarray(U, dtype=np.int32)
    V = np.asarray(V, dtype=np.int32)
    T = np.asarray(T, dtype=np.int32)
    E = U.size
    order = np.argsort(T, kind="stable")
    Us, Vs, Ts = U[order], V[order], T[order]
    ids = order.astype(np.int32)
    inter = np.empty(2 * E, dtype=np.int64)
    inter[0::2] = Us
    inter[1::2] = Vs
    uniq, first = np.unique(inter, return_index=True)
    nV = uniq.size
    new_of_uniq = np.empty(nV, dtype=np.int64)
    new_of_uniq[np.argsort(first, kind="stable")] = np.arange(nV)
    Un = new_of_uniq[np.searchsorted(uniq, Us)].astype(np.int32)
    Vn = new_of_uniq[np.searchsorted(uniq, Vs)].astype(np.int32)
    outV = np.argsort(Un, kind="stable").astype(np.int32)
    inV = np.argsort(Vn, kind="stable").astype(np.int32)
    outR = np.zeros(nV + 1, dtype=np.int32)
    outR[1:] = np.cumsum(np.bincount(Un, minlength=nV))
    inR = np.zeros(nV + 1, dtype=np.int32)
    inR[1:] = np.cumsum(np.bincount(Vn, minlength=nV))
    Eg = np.stack([Un, Vn, Ts], axis=1).astype(np.int32)
    return dict(E=E, V=nV, Eg=np.ascontiguousarray(Eg), outV=outV, outR=outR,
                inV=inV, inR=inR, Euid=ids, T=Ts.copy(), rename=uniq[np.argsort(first, kind="stable")])
