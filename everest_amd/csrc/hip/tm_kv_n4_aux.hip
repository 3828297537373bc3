// Instrumented, metered and enumeration variants, node map of 4 (see tm_variants.h).
#include "tm_variants.cuh"
TMV_DEF(4, 0, 1, 1, 0)
TMV_DEF(4, 1, 0, 1, 0)
TMV_DEF(4, 0, 2, 1, 0)
