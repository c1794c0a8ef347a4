// ORACLE — test infrastructure only (see oracle.h). Progressive alignment
// along a guide tree, the same merges as tests/oracle_bind.progressive_msa
// (profile counts, fp32 column scores of orc_profile_scores, the recurrence
// and traceback of orc_nw_flags_sc / orc_nw_path, gap columns inserted into
// every member), in C++ so that sets of hundreds or thousands of sequences
// and profiles of 20,000 columns finish in seconds. Never linked into the
// product.
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

namespace {

const char kLetters[] = "ARNDCQEGHILKMFPSTWYVBZX*";

int letter_index(char ch) {
  if (ch >= 'a' && ch <= 'z') ch = (char)(ch - 'a' + 'A');
  const char* p = std::strchr(kLetters, ch);
  return (p && ch) ? (int)(p - kLetters) : 22;
}

struct Group {
  std::vector<int> members;
  std::vector<std::string> rows;
};

// Column frequencies of a group in DP order (columns reversed), 24 letters;
// the gap count is not used by the column score.
std::vector<float> freqs(const Group& g) {
  const int L = (int)g.rows[0].size();
  const int k = (int)g.rows.size();
  std::vector<int32_t> cnt((size_t)L * 24, 0);
  for (const std::string& r : g.rows)
    for (int t = 0; t < L; ++t)
      if (r[t] != '-') ++cnt[(size_t)(L - 1 - t) * 24 + letter_index(r[t])];
  std::vector<float> f((size_t)L * 24);
  for (size_t q = 0; q < f.size(); ++q) f[q] = (float)cnt[q] / (float)k;
  return f;
}

// Merge path of X (columns) against Y (rows): 'D', 'L', 'U' in forward order.
std::string profile_path(const Group& gx, const Group& gy, const int8_t* S, int go, int ge) {
  const int n = (int)gx.rows[0].size(), m = (int)gy.rows[0].size();
  const std::vector<float> fX = freqs(gx), fY = freqs(gy);
  // P_r(a) = sum_b fY_r(b) * S[b][a], in b order (orc_profile_scores)
  std::vector<float> P((size_t)m * 24);
  for (int r = 0; r < m; ++r)
    for (int a = 0; a < 24; ++a) {
      float acc = 0.0f;
      for (int b = 0; b < 24; ++b) {
        float p = fY[(size_t)r * 24 + b] * (float)S[24 * b + a];
        acc = acc + p;
      }
      P[(size_t)r * 24 + a] = acc;
    }
  // X frequencies by letter, then column: a row's scores are one pass per
  // letter over the columns, each cell's sum still taken in letter order
  std::vector<float> fXT((size_t)24 * n);
  for (int c = 0; c < n; ++c)
    for (int a = 0; a < 24; ++a) fXT[(size_t)a * n + c] = fX[(size_t)c * 24 + a];
  std::vector<uint8_t> nib((size_t)n * m);
  std::vector<int32_t> Hrow(n), Frow(n);
  for (int c = 0; c < n; ++c) {
    Hrow[c] = -ge * (c + 1);
    Frow[c] = ORC_NEG_INF;
  }
  const int B = 64;  // rows of column scores computed at a time
  std::vector<int32_t> sc((size_t)B * n);
  for (int r0 = 0; r0 < m; r0 += B) {
    const int r1 = r0 + B < m ? r0 + B : m;
#pragma omp parallel
    {
      std::vector<float> acc(n);
#pragma omp for schedule(static)
      for (int r = r0; r < r1; ++r) {
        for (int c = 0; c < n; ++c) acc[c] = 0.0f;
        for (int a = 0; a < 24; ++a) {
          const float pa = P[(size_t)r * 24 + a];
          const float* fa = &fXT[(size_t)a * n];
          for (int c = 0; c < n; ++c) {
            float p = fa[c] * pa;
            acc[c] = acc[c] + p;
          }
        }
        int32_t* out = &sc[(size_t)(r - r0) * n];
        for (int c = 0; c < n; ++c) out[c] = (int32_t)__builtin_rintf(acc[c]);
      }
    }
    for (int r = r0; r < r1; ++r) {
      int32_t Hdiag = -ge * r, Hleft = -ge * (r + 1), Eleft = ORC_NEG_INF;
      const int32_t* srow = &sc[(size_t)(r - r0) * n];
      for (int c = 0; c < n; ++c) {
        const int32_t Fext = Frow[c] - ge;
        const int32_t F = Fext > Hrow[c] - go ? Fext : Hrow[c] - go;
        const int cu = F == Fext;
        const int32_t Eext = Eleft - ge;
        const int32_t E = Eext > Hleft - go ? Eext : Hleft - go;
        const int cl = E == Eext;
        const int32_t D = Hdiag + srow[c];
        int32_t H = E > F ? E : F;
        H = H > D ? H : D;
        const int left = ((H == E) && (H != F)) || (H == D);
        const int up = (H == F) || (H == D);
        nib[(size_t)r * n + c] = (uint8_t)((left << 3) | (up << 2) | (cu << 1) | cl);
        Hdiag = Hrow[c];
        Hrow[c] = H;
        Frow[c] = F;
        Hleft = H;
        Eleft = E;
      }
    }
  }
  std::string ops((size_t)n + m + 1, '\0');
  const int len = orc_nw_path(nib.data(), n, m, &ops[0]);
  ops.resize(len);
  return ops;
}

}  // namespace

extern "C" int orc_progressive_msa(const char* ascii, const int64_t* off, int n0,
                                   const int32_t* joins3, int njoins, const int8_t* S,
                                   int go, int ge, char* rows_out, int64_t stride,
                                   int* msa_len, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  std::map<int, Group> groups;
  for (int i = 0; i < n0; ++i) {
    Group g;
    g.members.push_back(i);
    std::string s;
    for (int64_t t = off[i]; t < off[i + 1]; ++t) s.push_back(kLetters[letter_index(ascii[t])]);
    g.rows.push_back(s);
    groups[i] = std::move(g);
  }
  for (int k = 0; k < njoins; ++k) {
    const int node = joins3[3 * k], l = joins3[3 * k + 1], r = joins3[3 * k + 2];
    auto il = groups.find(l), ir = groups.find(r);
    if (il == groups.end() || ir == groups.end() || groups.count(node)) return -1;
    Group gx = std::move(il->second), gy = std::move(ir->second);
    groups.erase(l);
    groups.erase(r);
    std::string ops;
    if (gx.rows[0].empty())
      ops.assign(gy.rows[0].size(), 'U');
    else if (gy.rows[0].empty())
      ops.assign(gx.rows[0].size(), 'L');
    else
      ops = profile_path(gx, gy, S, go, ge);
    Group g;
    g.members = gx.members;
    g.members.insert(g.members.end(), gy.members.begin(), gy.members.end());
    for (int side = 0; side < 2; ++side) {
      const Group& src = side ? gy : gx;
      const char skip = side ? 'L' : 'U';  // the op that puts a gap in this side
      for (const std::string& row : src.rows) {
        std::string out;
        out.reserve(ops.size());
        size_t ix = 0;
        for (char op : ops) out.push_back(op == skip ? '-' : row[ix++]);
        g.rows.push_back(std::move(out));
      }
    }
    groups[node] = std::move(g);
  }
  if (groups.size() != 1) return -1;
  const Group& g = groups.begin()->second;
  const int L = (int)g.rows[0].size();
  *msa_len = L;
  if (rows_out) {
    if (stride < L + 1) return -2;
    for (size_t q = 0; q < g.members.size(); ++q) {
      std::memcpy(rows_out + (int64_t)g.members[q] * stride, g.rows[q].data(), L);
      rows_out[(int64_t)g.members[q] * stride + L] = '\0';
    }
  }
  return 0;
}
