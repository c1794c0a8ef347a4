// Seeded synthetic protein sets of the benchmark workloads (host code).
//
//   0  tree64     64 sequences evolved from one random root of length 225 down
//                 a random binary tree (substitution 0.15, indel 0.01/site),
//                 lengths kept in [150, 300]
//   1  tree1000   1,000 sequences, root 400, same process, lengths in [200, 600]
//   2  lognormal  4,000 sequences, lengths log-normal (median 350, sigma 0.8)
//                 clipped to [50, 5000], each from its own random parent with
//                 substitution 0.3
//   3  long       256 sequences, parent lengths uniform in [8000, 20000],
//                 substitution 0.2, indel 0.02
//   4  job        one MSA job of the served stream: root length uniform in
//                 [100, 2000], tree process as 0, lengths kept in [100, 2000]
// Residues are uniform over the 20 standard amino acids; a substitution draws
// a uniform residue; an indel is an insertion or a deletion with probability
// 1/2, of geometric length with mean 3. The generator is std::mt19937_64 with
// distributions written out here, so every platform produces the same set.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mestrado_synth.h"

namespace {

const char kStd20[] = "ACDEFGHIKLMNPQRSTVWY";

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  uint64_t u64() { return g(); }
  double uni() { return (double)(u64() >> 11) * (1.0 / 9007199254740992.0); }
  int below(int n) { return (int)(u64() % (uint64_t)n); }
  int range(int lo, int hi) { return lo + below(hi - lo + 1); }  // inclusive
  char residue() { return kStd20[below(20)]; }
  int geometric_mean3() {  // support 1,2,...; P(k) = (2/3)^(k-1) (1/3)
    double u = uni();
    if (u <= 0.0) u = 1e-300;
    return 1 + (int)std::floor(std::log(u) / std::log(2.0 / 3.0));
  }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 <= 0.0) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

std::string random_seq(Rng& r, int len) {
  std::string s(len, 'A');
  for (int i = 0; i < len; ++i) s[i] = r.residue();
  return s;
}

std::string mutate(Rng& r, const std::string& p, double subst, double indel) {
  std::string out;
  out.reserve(p.size() + 16);
  size_t i = 0;
  while (i < p.size()) {
    if (indel > 0 && r.uni() < indel) {
      int L = r.geometric_mean3();
      if (r.uni() < 0.5) {
        for (int k = 0; k < L; ++k) out.push_back(r.residue());
      } else {
        i += L;
        continue;
      }
    }
    out.push_back(r.uni() < subst ? r.residue() : p[i]);
    ++i;
  }
  return out;
}

void tree_evolve(Rng& r, const std::string& node, int leaves, double subst, double indel,
                 int lo, int hi, std::vector<std::string>& out) {
  if (leaves == 1) {
    out.push_back(node);
    return;
  }
  int k1 = r.range(1, leaves - 1);
  int k[2] = {k1, leaves - k1};
  for (int side = 0; side < 2; ++side) {
    std::string child;
    for (int tries = 0;; ++tries) {
      child = mutate(r, node, subst, indel);
      if ((int)child.size() >= lo && (int)child.size() <= hi) break;
      if (tries == 100) {  // keep the range: trim or extend
        while ((int)child.size() > hi) child.pop_back();
        while ((int)child.size() < lo) child.push_back(r.residue());
        break;
      }
    }
    tree_evolve(r, child, k[side], subst, indel, lo, hi, out);
  }
}

std::vector<std::string> make_set(int config, uint64_t seed, int n) {
  Rng r(seed * 0x9E3779B97F4A7C15ull + (uint64_t)config * 1000003ull + 12345);
  std::vector<std::string> out;
  switch (config) {
    case 0:
    case 1: {
      int N = n > 0 ? n : (config == 0 ? 64 : 1000);
      int root = config == 0 ? 225 : 400;
      int lo = config == 0 ? 150 : 200, hi = config == 0 ? 300 : 600;
      tree_evolve(r, random_seq(r, root), N, 0.15, 0.01, lo, hi, out);
      break;
    }
    case 2: {
      int N = n > 0 ? n : 4000;
      for (int i = 0; i < N; ++i) {
        double L = std::exp(std::log(350.0) + 0.8 * r.normal());
        int len = (int)std::lround(L);
        len = len < 50 ? 50 : (len > 5000 ? 5000 : len);
        out.push_back(mutate(r, random_seq(r, len), 0.3, 0.0));
      }
      break;
    }
    case 3: {
      int N = n > 0 ? n : 256;
      for (int i = 0; i < N; ++i)
        out.push_back(mutate(r, random_seq(r, r.range(8000, 20000)), 0.2, 0.02));
      break;
    }
    case 4: {
      // job size from the caller (n); power-law sizes are drawn by mst_synth_job_sizes
      int N = n > 0 ? n : 8;
      tree_evolve(r, random_seq(r, r.range(100, 2000)), N, 0.15, 0.01, 100, 2000, out);
      break;
    }
    default:
      break;
  }
  return out;
}

}  // namespace

extern "C" int mst_synth(int config, uint64_t seed, int n, char* ascii, int64_t cap,
                         int64_t* offsets, int64_t offsets_cap, int* n_out, int64_t* total_out) {
  if (config < 0 || config > 4) return -1;
  std::vector<std::string> s = make_set(config, seed, n);
  int64_t total = 0;
  for (auto& x : s) total += (int64_t)x.size();
  if (n_out) *n_out = (int)s.size();
  if (total_out) *total_out = total;
  if (!ascii || !offsets) return 0;
  if (cap < total || offsets_cap < (int64_t)s.size() + 1) return -2;
  int64_t o = 0;
  offsets[0] = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    memcpy(ascii + o, s[i].data(), s[i].size());
    o += (int64_t)s[i].size();
    offsets[i + 1] = o;
  }
  return 0;
}

extern "C" int mst_synth_job_sizes(uint64_t seed, int njobs, int lo, int hi, double exponent,
                                   int* sizes) {
  // Power law p(N) ~ N^-exponent on [lo, hi], sampled by inverting the
  // continuous distribution and rounding down.
  if (njobs < 0 || lo < 1 || hi < lo || !sizes) return -1;
  Rng r(seed ^ 0xA5A5A5A5DEADBEEFull);
  double a = 1.0 - exponent;
  double L = std::pow((double)lo, a), H = std::pow((double)hi + 1.0, a);
  for (int k = 0; k < njobs; ++k) {
    double u = r.uni();
    int v = (int)std::floor(std::pow(L + u * (H - L), 1.0 / a));
    sizes[k] = v < lo ? lo : (v > hi ? hi : v);
  }
  return 0;
}
