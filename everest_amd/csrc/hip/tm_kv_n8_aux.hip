// Instrumented, metered and enumeration variants, node map of 8 (see tm_variants.h).
#include "tm_variants.cuh"
TMV_DEF(8, 0, 1, 1, 0)
TMV_DEF(8, 1, 0, 1, 0)
TMV_DEF(8, 0, 2, 1, 0)
