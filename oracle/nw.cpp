// ORACLE — test infrastructure only (see oracle.h). CPU restatement of the
// reference's Needleman-Wunsch recurrences; never linked into the product.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Letter order of the substitution matrix (definitions.cuh:96-121 rows,
// revConvConst definitions.cuh:134-137).
const char kLetters[] = "ARNDCQEGHILKMFPSTWYVBZX*";

struct EncTable {
  int8_t t[256];
  EncTable() {
    for (int i = 0; i < 256; ++i) t[i] = 22;  // unknown -> X (build rule D8)
    for (int k = 0; k < 24; ++k) {
      unsigned char c = (unsigned char)kLetters[k];
      t[c] = (int8_t)k;
      if (c >= 'A' && c <= 'Z') t[c - 'A' + 'a'] = (int8_t)k;
    }
  }
};
const EncTable kEnc;

inline int32_t mx(int32_t a, int32_t b) { return a > b ? a : b; }

}  // namespace

extern "C" int orc_encode_reverse(const char* ascii, int n, int8_t* out) {
  // AlignmentScore.cu:168-177: the sequence is reversed while encoding.
  for (int i = 0; i < n; ++i) out[i] = kEnc.t[(unsigned char)ascii[n - 1 - i]];
  return n;
}

extern "C" char orc_decode(int code) {
  return (code >= 0 && code < 24) ? kLetters[code] : '?';
}

// R1: NeedlemanWunsch.cu:22-161. X indexes columns, Y rows; S is read as
// S[Y][X] (:140). Row -1: H=-ge(c+1), F=-inf (:65-70). Column -1: H=-ge(r+1),
// E=-inf (:94-99). Corner H(-1,-1)=0 (:89).
extern "C" int32_t orc_nw_score(const int8_t* X, int n, const int8_t* Y, int m,
                                const int8_t* S, int go, int ge) {
  if (n == 0) return -ge * m;
  if (m == 0) return -ge * n;
  std::vector<int32_t> Hrow(n), Frow(n);
  for (int c = 0; c < n; ++c) {
    Hrow[c] = -ge * (c + 1);
    Frow[c] = ORC_NEG_INF;
  }
  int32_t H = 0;
  for (int r = 0; r < m; ++r) {
    int32_t Hdiag = -ge * r;        // H(r-1,-1)
    int32_t Hleft = -ge * (r + 1);  // H(r,-1)
    int32_t Eleft = ORC_NEG_INF;
    const int8_t* Srow = S + 24 * Y[r];
    for (int c = 0; c < n; ++c) {
      int32_t F = mx(Frow[c] - ge, Hrow[c] - go);  // :142
      int32_t E = mx(Eleft - ge, Hleft - go);      // :143
      H = mx(mx(E, F), Hdiag + Srow[X[c]]);        // :145-146
      Hdiag = Hrow[c];
      Hrow[c] = H;
      Frow[c] = F;
      Hleft = H;
      Eleft = E;
    }
  }
  return H;
}

// R2: NeedlemanWunsch.cu:310-347. Same recurrence plus one nibble per cell.
extern "C" int32_t orc_nw_flags(const int8_t* X, int n, const int8_t* Y, int m,
                                const int8_t* S, int go, int ge, uint8_t* nib) {
  if (n == 0) return -ge * m;
  if (m == 0) return -ge * n;
  std::vector<int32_t> Hrow(n), Frow(n);
  for (int c = 0; c < n; ++c) {
    Hrow[c] = -ge * (c + 1);
    Frow[c] = ORC_NEG_INF;
  }
  int32_t H = 0;
  for (int r = 0; r < m; ++r) {
    int32_t Hdiag = -ge * r;
    int32_t Hleft = -ge * (r + 1);
    int32_t Eleft = ORC_NEG_INF;
    const int8_t* Srow = S + 24 * Y[r];
    for (int c = 0; c < n; ++c) {
      int32_t Fext = Frow[c] - ge;
      int32_t F = mx(Fext, Hrow[c] - go);
      int cu = (F == Fext);                         // :316
      int32_t Eext = Eleft - ge;
      int32_t E = mx(Eext, Hleft - go);
      int cl = (E == Eext);                         // :336
      int32_t D = Hdiag + Srow[X[c]];
      H = mx(mx(E, F), D);
      int left = ((H == E) && (H != F)) || (H == D);  // :330
      int up = (H == F) || (H == D);                  // :332
      nib[(size_t)r * n + c] = (uint8_t)((left << 3) | (up << 2) | (cu << 1) | cl);
      Hdiag = Hrow[c];
      Hrow[c] = H;
      Frow[c] = F;
      Hleft = H;
      Eleft = E;
    }
  }
  return H;
}

namespace {
// Boundary nibbles of NeedlemanWunsch.cu:233,238,240: row -1 = LEFT+contLeft,
// column -1 = UP+contUp, corner = STOP.
inline int nib_at(const uint8_t* nib, int n, int x, int y) {
  if (x < 0 && y < 0) return 0;
  if (y < 0) return 9;
  if (x < 0) return 6;
  return nib[(size_t)y * n + x];
}
enum { kUp = 4, kLeft = 8, kCross = 12, kDir = 12, kContUp = 2, kContLeft = 1 };

// Walks the traceback of R3 (NeedlemanWunsch.cu:409-482) and calls emit(todo,x,y)
// for each move taken from cell (x,y). Going UP into row -1 continues along
// row -1 (build fix D9) instead of stopping.
template <class F>
int walk(const uint8_t* nib, int n, int m, F emit) {
  int x = n - 1, y = m - 1;
  int prev = kCross;
  int cur = nib_at(nib, n, x, y);
  int steps = 0;
  while (cur & kDir) {
    int todo;
    if ((prev & kDir) == kUp && (prev & kContUp))
      todo = kUp;
    else if ((prev & kDir) == kLeft && (prev & kContLeft))
      todo = kLeft;
    else if ((cur & kDir) == kUp)
      todo = kUp;
    else if ((cur & kDir) == kLeft)
      todo = kLeft;
    else
      todo = kCross;
    emit(todo, x, y);
    if (todo == kLeft)
      --x;
    else if (todo == kUp)
      --y;
    else {
      --x;
      --y;
    }
    prev = todo | (cur & 3);
    cur = nib_at(nib, n, x, y);
    ++steps;
  }
  return steps;
}
}  // namespace

extern "C" int orc_nw_traceback(const uint8_t* nib, const int8_t* X, int n,
                                const int8_t* Y, int m, char* outX, char* outY) {
  int k = 0;
  walk(nib, n, m, [&](int todo, int x, int y) {
    if (todo == kLeft) {
      outX[k] = orc_decode(X[x]);
      outY[k] = '-';
    } else if (todo == kUp) {
      outX[k] = '-';
      outY[k] = orc_decode(Y[y]);
    } else {
      outX[k] = orc_decode(X[x]);
      outY[k] = orc_decode(Y[y]);
    }
    ++k;
  });
  outX[k] = 0;
  outY[k] = 0;
  return k;
}

extern "C" int orc_nw_path(const uint8_t* nib, int n, int m, char* ops) {
  int k = 0;
  walk(nib, n, m, [&](int todo, int, int) {
    ops[k++] = todo == kLeft ? 'L' : (todo == kUp ? 'U' : 'D');
  });
  ops[k] = 0;
  return k;
}

// Textbook Gotoh over prefix lengths (i rows of Y, j columns of X) in int64
// with a true minus infinity. Boundary gaps along the DP start cost ge per
// residue, every other gap opens with go.
extern "C" int64_t orc_gotoh_score(const int8_t* X, int n, const int8_t* Y,
                                   int m, const int8_t* S, int go, int ge) {
  const int64_t NI = INT64_MIN / 4;
  std::vector<int64_t> V((size_t)(m + 1) * (n + 1)), E(V.size()), F(V.size());
  auto at = [&](int i, int j) { return (size_t)i * (n + 1) + j; };
  V[at(0, 0)] = 0;
  E[at(0, 0)] = F[at(0, 0)] = NI;
  for (int j = 1; j <= n; ++j) {
    V[at(0, j)] = -(int64_t)ge * j;
    E[at(0, j)] = NI;
    F[at(0, j)] = NI;
  }
  for (int i = 1; i <= m; ++i) {
    V[at(i, 0)] = -(int64_t)ge * i;
    E[at(i, 0)] = NI;
    F[at(i, 0)] = NI;
    for (int j = 1; j <= n; ++j) {
      int64_t Mv = V[at(i - 1, j - 1)] + S[24 * Y[i - 1] + X[j - 1]];
      int64_t Ev = std::max(E[at(i, j - 1)] - ge, V[at(i, j - 1)] - go);
      int64_t Fv = std::max(F[at(i - 1, j)] - ge, V[at(i - 1, j)] - go);
      E[at(i, j)] = Ev;
      F[at(i, j)] = Fv;
      V[at(i, j)] = std::max(Mv, std::max(Ev, Fv));
    }
  }
  return V[at(m, n)];
}

// Cost model stated in words: walk the alignment columns in DP order (the
// reverse of the forward strings); a residue pair scores S; a maximal run of
// one gap type costs ge per column if it is the first run of the alignment,
// else go for its first column and ge for each further column.
extern "C" int64_t orc_rescore(const char* gx, const char* gy, int len,
                               const int8_t* S, int go, int ge) {
  int64_t s = 0;
  int prevType = -1;  // 0 pair, 1 gap in Y (X residue), 2 gap in X
  bool firstRun = true;
  for (int t = len - 1; t >= 0; --t) {
    char a = gx[t], b = gy[t];
    int type = (a != '-' && b != '-') ? 0 : (b == '-' ? 1 : 2);
    if (prevType != -1 && type != prevType) firstRun = false;
    if (type == 0)
      s += S[24 * kEnc.t[(unsigned char)b] + kEnc.t[(unsigned char)a]];
    else if (type == prevType)
      s -= ge;
    else
      s -= firstRun ? ge : go;
    prevType = type;
  }
  return s;
}

namespace {
struct Brute {
  const int8_t *X, *Y, *S;
  int n, m, go, ge;
  int64_t best;
  std::string gx, gy;
  void rec(int i, int j) {  // i: X consumed, j: Y consumed (DP order)
    if (i == n && j == m) {
      std::string fx(gx.rbegin(), gx.rend()), fy(gy.rbegin(), gy.rend());
      int64_t s = orc_rescore(fx.c_str(), fy.c_str(), (int)fx.size(), S, go, ge);
      if (s > best) best = s;
      return;
    }
    if (i < n && j < m) {
      gx.push_back(orc_decode(X[i]));
      gy.push_back(orc_decode(Y[j]));
      rec(i + 1, j + 1);
      gx.pop_back();
      gy.pop_back();
    }
    if (i < n) {
      gx.push_back(orc_decode(X[i]));
      gy.push_back('-');
      rec(i + 1, j);
      gx.pop_back();
      gy.pop_back();
    }
    if (j < m) {
      gx.push_back('-');
      gy.push_back(orc_decode(Y[j]));
      rec(i, j + 1);
      gx.pop_back();
      gy.pop_back();
    }
  }
};
}  // namespace

extern "C" int64_t orc_bruteforce(const int8_t* X, int n, const int8_t* Y,
                                  int m, const int8_t* S, int go, int ge) {
  Brute b{X, Y, S, n, m, go, ge, INT64_MIN, {}, {}};
  b.rec(0, 0);
  return b.best;
}

extern "C" double orc_pair_list(const int8_t* seqs, const int64_t* off,
                                const int32_t* pi, const int32_t* pj,
                                int64_t npairs, const int8_t* S, int go, int ge,
                                int32_t* out, int nthreads) {
  // Longest pairs first (dynamic schedule) for load balance.
  std::vector<int64_t> order(npairs);
  for (int64_t k = 0; k < npairs; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    int64_t ca = (off[pi[a] + 1] - off[pi[a]]) * (off[pj[a] + 1] - off[pj[a]]);
    int64_t cb = (off[pi[b] + 1] - off[pi[b]]) * (off[pj[b] + 1] - off[pj[b]]);
    return ca > cb;
  });
  auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int64_t t = 0; t < npairs; ++t) {
    int64_t k = order[t];
    int i = pi[k], j = pj[k];
    out[k] = orc_nw_score(seqs + off[i], (int)(off[i + 1] - off[i]), seqs + off[j],
                          (int)(off[j + 1] - off[j]), S, go, ge);
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

extern "C" double orc_all_pairs(const int8_t* seqs, const int64_t* off, int n,
                                const int8_t* S, int go, int ge, int32_t* out,
                                int nthreads, int64_t pair_limit) {
  std::vector<int32_t> pi, pj;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      if (pair_limit > 0 && (int64_t)pi.size() >= pair_limit) break;
      pi.push_back(i);
      pj.push_back(j);
    }
  std::vector<int32_t> sc(pi.size());
  double secs = orc_pair_list(seqs, off, pi.data(), pj.data(), (int64_t)pi.size(),
                              S, go, ge, sc.data(), nthreads);
  for (size_t k = 0; k < pi.size(); ++k) {
    out[(size_t)pi[k] * n + pj[k]] = sc[k];
    out[(size_t)pj[k] * n + pi[k]] = sc[k];
  }
  return secs;
}

// ---- profile-profile DP (build-defined; reduces to R1-R3 on one-hot) ----
// Column frequencies f = count/k in fp32. P_j(a) = sum_b fY_j(b)*S[b][a] in b
// order, then s(i,j) = sum_a fX_i(a)*P_j(a) in a order, every product and sum
// rounded to fp32 separately, finally rounded half-to-even to int32.
extern "C" void orc_profile_scores(const int32_t* cntX, int n, int kX,
                                   const int32_t* cntY, int m, int kY,
                                   const int8_t* S, int32_t* sc) {
  std::vector<float> fX((size_t)n * 24), P((size_t)m * 24);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 24; ++a) fX[(size_t)i * 24 + a] = (float)cntX[i * 25 + a] / (float)kX;
  for (int j = 0; j < m; ++j) {
    float fy[24];
    for (int b = 0; b < 24; ++b) fy[b] = (float)cntY[j * 25 + b] / (float)kY;
    for (int a = 0; a < 24; ++a) {
      float acc = 0.0f;
      for (int b = 0; b < 24; ++b) {
        float p = fy[b] * (float)S[24 * b + a];
        acc = acc + p;
      }
      P[(size_t)j * 24 + a] = acc;
    }
  }
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < n; ++i) {
      float acc = 0.0f;
      for (int a = 0; a < 24; ++a) {
        float p = fX[(size_t)i * 24 + a] * P[(size_t)j * 24 + a];
        acc = acc + p;
      }
      sc[(size_t)j * n + i] = (int32_t)__builtin_rintf(acc);
    }
}

extern "C" int32_t orc_nw_flags_sc(const int32_t* sc, int n, int m, int go,
                                   int ge, uint8_t* nib) {
  if (n == 0) return -ge * m;
  if (m == 0) return -ge * n;
  std::vector<int32_t> Hrow(n), Frow(n);
  for (int c = 0; c < n; ++c) {
    Hrow[c] = -ge * (c + 1);
    Frow[c] = ORC_NEG_INF;
  }
  int32_t H = 0;
  for (int r = 0; r < m; ++r) {
    int32_t Hdiag = -ge * r;
    int32_t Hleft = -ge * (r + 1);
    int32_t Eleft = ORC_NEG_INF;
    for (int c = 0; c < n; ++c) {
      int32_t Fext = Frow[c] - ge;
      int32_t F = mx(Fext, Hrow[c] - go);
      int cu = (F == Fext);
      int32_t Eext = Eleft - ge;
      int32_t E = mx(Eext, Hleft - go);
      int cl = (E == Eext);
      int32_t D = Hdiag + sc[(size_t)r * n + c];
      H = mx(mx(E, F), D);
      int left = ((H == E) && (H != F)) || (H == D);
      int up = (H == F) || (H == D);
      nib[(size_t)r * n + c] = (uint8_t)((left << 3) | (up << 2) | (cu << 1) | cl);
      Hdiag = Hrow[c];
      Hrow[c] = H;
      Frow[c] = F;
      Hleft = H;
      Eleft = E;
    }
  }
  return H;
}
