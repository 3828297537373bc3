// Multi-query (batch) counting variants, node map of 4 (see tm_variants.h).
#include "tm_variants.cuh"
TMV_DEF(4, 0, 0, 1, 1)
TMV_DEF(4, 0, 0, 4, 1)
TMV_DEF(4, 0, 0, 5, 1)
TMV_DEF(4, 0, 0, 6, 1)
