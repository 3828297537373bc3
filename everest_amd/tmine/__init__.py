"""Query frontend of the temporal-motif miner (Python side of the query path).

Mirrors the reference package system/tmine: Query (YAML queries), KernelGen
(per-query worker source, here HIP for gfx950 from templates/), frontend
(Mining / MiningSingle), locations and utils.
"""
