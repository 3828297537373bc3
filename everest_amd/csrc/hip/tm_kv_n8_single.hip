// Single-query counting variants, node map of 8 (see tm_variants.h).
#include "tm_variants.cuh"
TMV_DEF(8, 0, 0, 1, 0)
TMV_DEF(8, 0, 0, 4, 0)
TMV_DEF(8, 0, 0, 5, 0)
TMV_DEF(8, 0, 0, 6, 0)
