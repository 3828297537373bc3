// everest/gen.h -- deterministic synthetic temporal graphs for the benchmark
// workloads (no reference counterpart: the reference reads SNAP downloads).
// Outputs are U, V, T arrays of E ints in generation (= line) order.
#pragma once
#include <cstddef>
#include <cstdint>

namespace everest::gen {

// Endpoints Zipf(alpha) over ranks 1..N (node id = rank - 1), v != u,
// times uniform integers in [0, span).
void zipf_uniform(size_t N, size_t E, double alpha, int64_t span, uint64_t seed, int *U, int *V, int *T);

// Chung-Lu power-law graph with per-source bursty Poisson times over [0, span).
void powerlaw_bursty(size_t N, size_t E, double alpha, double sigma, int64_t span, uint64_t seed, int *U, int *V,
                     int *T);

}  // namespace everest::gen
