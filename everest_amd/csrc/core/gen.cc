// gen.cc -- deterministic synthetic temporal graphs (host, multithreaded).
//
// Random numbers come from a counter-based generator: draw i of stream s is
// splitmix64(seed, s, i) mapped to a double in [0, 1).  Every value depends
// only on (seed, stream, index), so the output is identical for any thread
// count.  Edges are produced in generation order, which is the edge-file line
// order and fixes the order of equal timestamps after the stable time sort.
//
// zipf_uniform: both endpoints Zipf(alpha) over ranks 1..N (P(r) ~ r^-alpha,
//   node id = r - 1), v redrawn while v == u, times uniform over `span`.
// powerlaw_bursty: Chung-Lu power-law graph.  Node out/in weights are
//   independent draws w = floor((1-u)^(-1/(alpha-1))) (a discretised Pareto,
//   P(W >= k) = k^-(alpha-1)), capped at the natural cutoff N^(1/(alpha-1));
//   u ~ w_out, v ~ w_in (alias tables), self-loops
//   redrawn, repeated (u, v) pairs kept.  Every source runs a Poisson process
//   of rate lambda_u ~ LogNormal(0, sigma) events/day from a uniform random
//   phase, wrapped at `span`; its out-edges take the process's event times in
//   generation order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <vector>

#include "everest/gen.h"

namespace everest::gen {

namespace {

inline uint64_t mix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

inline double unif(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t x = mix(mix(seed * 0x100000001b3ull + stream) ^ (i * 0xd1b54a32d192ed03ull));
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

template <class F>
void parallel_for(size_t n, F f) {
  int nt = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 100000) nt = 1;
  std::vector<std::thread> ts;
  for (int t = 0; t < nt; t++) {
    size_t b = n * t / nt, e = n * (t + 1) / nt;
    ts.emplace_back([=] {
      for (size_t i = b; i < e; i++) f(i);
    });
  }
  for (auto &t : ts) t.join();
}

// Vose alias table over non-negative weights.
struct Alias {
  std::vector<double> prob;
  std::vector<int> alias;
  explicit Alias(const std::vector<double> &w) {
    size_t n = w.size();
    prob.assign(n, 0.0);
    alias.assign(n, 0);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> p(n);
    std::vector<int> small, large;
    for (size_t i = 0; i < n; i++) {
      p[i] = w[i] * (double)n / sum;
      (p[i] < 1.0 ? small : large).push_back((int)i);
    }
    while (!small.empty() && !large.empty()) {
      int s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = p[s];
      alias[s] = l;
      p[l] = (p[l] + p[s]) - 1.0;
      if (p[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (int l : large) prob[l] = 1.0;
    for (int s : small) prob[s] = 1.0;
  }
  int sample(double u1, double u2) const {
    size_t n = prob.size();
    size_t k = std::min(n - 1, (size_t)(u1 * (double)n));
    return u2 < prob[k] ? (int)k : alias[k];
  }
};

constexpr double DAY = 86400.0;

}  // namespace

void zipf_uniform(size_t N, size_t E, double alpha, int64_t span, uint64_t seed, int *U, int *V, int *T) {
  if (N < 2 || span < 1) throw std::runtime_error("zipf_uniform: need N >= 2 and span >= 1");
  std::vector<double> w(N);
  for (size_t r = 0; r < N; r++) w[r] = std::pow((double)(r + 1), -alpha);
  Alias a(w);
  parallel_for(E, [&](size_t i) {
    int u = a.sample(unif(seed, 1, i), unif(seed, 2, i));
    int v = u;
    for (uint64_t k = 0; v == u; k++) v = a.sample(unif(seed, 3, i + k * E), unif(seed, 4, i + k * E));
    U[i] = u;
    V[i] = v;
    int64_t t = (int64_t)(unif(seed, 5, i) * (double)span);
    T[i] = (int)std::min<int64_t>(t, span - 1);
  });
}

void powerlaw_bursty(size_t N, size_t E, double alpha, double sigma, int64_t span, uint64_t seed, int *U, int *V,
                     int *T) {
  if (N < 2 || span < 1 || alpha <= 1.0) throw std::runtime_error("powerlaw_bursty: bad parameters");
  std::vector<double> wout(N), win(N), lam(N), phase(N);
  // natural cutoff of a power-law sample: the expected largest of N draws
  const double cap = std::floor(std::pow((double)N, 1.0 / (alpha - 1.0)));
  parallel_for(N, [&](size_t k) {
    wout[k] = std::min(cap, std::floor(std::pow(1.0 - unif(seed, 11, k), -1.0 / (alpha - 1.0))));
    win[k] = std::min(cap, std::floor(std::pow(1.0 - unif(seed, 12, k), -1.0 / (alpha - 1.0))));
    double u1 = unif(seed, 13, k), u2 = unif(seed, 14, k);
    double z = std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * M_PI * u2);
    lam[k] = std::exp(sigma * z);
    phase[k] = unif(seed, 15, k) * (double)span;
  });
  Alias ao(wout), ai(win);
  parallel_for(E, [&](size_t i) {
    int u = ao.sample(unif(seed, 21, i), unif(seed, 22, i));
    int v = u;
    for (uint64_t k = 0; v == u; k++) v = ai.sample(unif(seed, 23, i + k * E), unif(seed, 24, i + k * E));
    U[i] = u;
    V[i] = v;
  });
  // group edges by source in generation order (counting sort)
  std::vector<int64_t> off(N + 1, 0);
  for (size_t i = 0; i < E; i++) off[(size_t)U[i] + 1]++;
  for (size_t k = 0; k < N; k++) off[k + 1] += off[k];
  std::vector<int> order(E);
  {
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (size_t i = 0; i < E; i++) order[fill[(size_t)U[i]]++] = (int)i;
  }
  const double sp = (double)span;
  parallel_for(N, [&](size_t k) {
    double acc = phase[k];
    const double scale = DAY / lam[k];
    for (int64_t s = off[k]; s < off[k + 1]; s++) {
      int i = order[(size_t)s];
      acc += -std::log1p(-unif(seed, 31, (uint64_t)i)) * scale;
      double t = std::fmod(acc, sp);
      T[i] = (int)std::min<int64_t>((int64_t)std::floor(t), span - 1);
    }
  });
}

}  // namespace everest::gen
