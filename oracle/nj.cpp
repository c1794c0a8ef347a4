// ORACLE — test infrastructure only (see oracle.h). Single-process restatement
// of the reference's neighbour joining (SpeciesBlock::buildTree and helpers),
// with the bins, priority queues, candidate list and q prediction kept, so the
// order in which pairs are examined and the q formula each pair is scored with
// follow the reference. Never linked into the product.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <queue>
#include <string>
#include <vector>

namespace {

// structs.h:163-181,183-211,213-235,237-263
struct Dij {
  float d;
  int i, j;
};
struct QDij {
  float d = FLT_MAX;
  int i = -1, j = -1;
  double q = DBL_MAX;
  void assign(double qq, float dd, int ii, int jj) {
    q = qq;
    d = dd;
    i = ii;
    j = jj;
  }
};
struct ActiveDij {
  float d;
  int i, j;
  bool active;
};
struct DijCmp {
  bool operator()(const Dij& a, const Dij& b) const { return a.d > b.d; }
};
struct QCmp {
  bool operator()(const QDij& a, const QDij& b) const { return a.q < b.q; }
};
typedef std::priority_queue<Dij, std::vector<Dij>, DijCmp> PQ;

struct NJ {
  // configuration (CmdLineOptions.cpp:79-88, SpeciesBlock.cpp:64-95)
  int n0, nn, numBins, binT, QDA;
  float ratio, pqRatio;
  bool predictQ;
  // state (SpeciesBlock.h:60-121) for one process
  int m_n, nextInternal, firstActive;
  std::vector<std::vector<float>> D;  // storage rows
  std::vector<double> T;
  std::vector<int> next, prev, redirect, binLoc;
  std::vector<double> binRange;
  std::vector<std::vector<int>> idxA;
  std::vector<std::pair<int, int>> binPair;
  std::vector<PQ> pqA;
  std::vector<int> binOrder;
  std::vector<float> vd;
  int rebuilds = 0, pqTriggers = 0;

  void init(const float* Din, int n) {
    n0 = m_n = n;
    nn = 2 * n - 1;
    binT = numBins * (numBins + 1) / 2;
    QDA = std::min(256, std::max(2, QDA));
    D.assign(n, std::vector<float>(n));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) D[i][j] = Din[(size_t)i * n + j];
    T.assign(n, 0.0);
    vd.assign(n, 0.0f);
    next.resize(nn);
    prev.resize(nn);
    redirect.assign(nn, -1);
    for (int i = 0; i < nn; ++i) {
      next[i] = i + 1;
      prev[i] = i - 1;
    }
    for (int i = 0; i < n; ++i) redirect[i] = i;
    nextInternal = n;
    firstActive = 0;
    binLoc.assign(n, -1);
    binRange.assign(numBins, 0.0);
    idxA.assign(numBins, std::vector<int>(numBins));
    int g = 0;
    for (int i = 0; i < numBins; ++i)
      for (int j = i; j < numBins; ++j) {
        idxA[i][j] = g;
        idxA[j][i] = g++;
      }
    binPair.resize(binT);
    g = 0;
    for (int i = 0; i < numBins; ++i)
      for (int j = i; j < numBins; ++j) binPair[g++] = {i, j};
    pqA.resize(binT);
    binOrder.resize(binT);
  }

  // SpeciesBlock.cpp:426-453
  void buildRowSums() {
    T.assign(n0, 0.0);
    for (int i = firstActive; i < nextInternal; i = next[i]) {
      int ri = redirect[i];
      double sum = 0.0;
      for (int j = firstActive; j < nextInternal; j = next[j]) sum += D[ri][redirect[j]];
      T[ri] = sum;
    }
  }

  // SpeciesBlock.cpp:456-674 (one process: every bin is local, lbin == gbin)
  void buildBins() {
    for (auto& q : pqA) q = PQ();
    double Tmax = -DBL_MAX;
    std::vector<double> Ts;
    for (int i = firstActive; i < nextInternal; i = next[i]) {
      double t = T[redirect[i]];
      Ts.push_back(t);
      if (t > Tmax) Tmax = t;
    }
    int rem = m_n % numBins, rdel = m_n / numBins, ir = rdel;
    std::sort(Ts.begin(), Ts.end());
    binRange[numBins - 1] = Tmax;
    for (int b = 0; b < numBins - 1; ++b) {
      int irNext = std::min(ir + 1, m_n - 1);
      // The reference reads past the end of its T vector when m_n < numBins - 1
      // (undefined); both indexes are clamped here.
      int a = std::min(ir, m_n - 1);
      binRange[b] = 0.5 * (Ts[a] + Ts[irNext]);
      ir += rdel + (rem > b);
    }
    binLoc.assign(n0, -1);
    for (int i = firstActive; i < nextInternal; i = next[i]) {
      int ri = redirect[i];
      int lb = (int)(std::lower_bound(binRange.begin(), binRange.end(), T[ri]) - binRange.begin());
      binLoc[ri] = std::min(lb, numBins - 1);
    }
    // Pairs are pushed per bin in row order (:544-562) and popped into the
    // heaps in receive order (:631-648); with one process each bin's heap
    // receives its own send buffer in push order.
    std::vector<std::vector<Dij>> send(binT);
    for (int i = firstActive; i < nextInternal; i = next[i]) {
      int ri = redirect[i];
      for (int j = next[i]; j < nextInternal; j = next[j]) {
        int rj = redirect[j];
        send[idxA[binLoc[ri]][binLoc[rj]]].push_back(Dij{D[ri][rj], i, j});
      }
    }
    for (int b = 0; b < binT; ++b)
      for (const Dij& x : send[b]) pqA[b].push(x);
    std::multimap<double, int, std::greater<double>> mm;
    for (int b = 0; b < binT; ++b)
      mm.insert(std::make_pair(binRange[binPair[b].first] + binRange[binPair[b].second], b));
    int c = 0;
    for (auto& kv : mm) binOrder[c++] = kv.second;
  }

  // SpeciesBlock.cpp:710-842
  void modifyMatrix(const QDij& minQ, int riMin) {
    auto unlink = [&](int x) {
      int p = prev[x], nx = next[x];
      prev[nx] = p;
      if (p == -1)
        firstActive = nx;
      else
        next[p] = nx;
    };
    unlink(minQ.j);
    unlink(minQ.i);
    std::vector<float> vdJ(n0), vdI(n0);
    int rj0 = redirect[minQ.j], ri0 = redirect[minQ.i];
    for (int k = firstActive; k < nextInternal; k = next[k]) {
      int rk = redirect[k];
      vdJ[rk] = D[rj0][rk];
      vdI[rk] = D[ri0][rk];
    }
    redirect[minQ.j] = redirect[minQ.i] = -1;
    double sum = 0.0;
    for (int k = firstActive; k < nextInternal; k = next[k]) {
      int rk = redirect[k];
      float Dik = vdI[rk], Djk = vdJ[rk];
      float tmp = ((Dik + Djk) - minQ.d) / 2;
      sum += tmp;
      vd[rk] = tmp;
      T[rk] += (double)(float)(tmp - (Dik + Djk));
    }
    T[riMin] = sum;
    for (int k = firstActive; k < nextInternal; k = next[k]) {
      int rk = redirect[k];
      D[rk][riMin] = vd[rk];
      D[riMin][rk] = vd[rk];
    }
  }

  static double branch(double minD, double Ti, double Tj, int n, double* R) {
    // TreeNode.cpp:30-58
    double L, Rr, half = 0.5;
    if (n == 2)
      Rr = L = (minD * half);
    else {
      L = half * (minD + (Ti - Tj) / (n - 2));
      Rr = half * (minD + (Tj - Ti) / (n - 2));
    }
    if (L < 0.0) {
      Rr -= L;
      L = 0.0;
    } else if (Rr < 0.0) {
      L -= Rr;
      Rr = 0.0;
    }
    *R = Rr;
    return L;
  }

  // SpeciesBlock.cpp:844-1343 for one process (the duplicate-removal pass is
  // dead code there because its zero-distance list is empty on entry).
  void run(orc_join_t* joins) {
    const int QDASZ = 512, pqRatioInit = 500, candidateScan = 50, minRebuild = 200;
    int qdAg_sz = QDA;
    std::vector<ActiveDij> adijA;
    std::vector<double> maxT2(numBins);
    std::vector<QDij> qdA(QDASZ), qdAg(qdAg_sz);
    int stepsUntilRebuild = std::max((int)(m_n * ratio), minRebuild);
    int n = 0;
    buildRowSums();
    buildBins();
    int pq_ratio_ck = pqRatioInit;
    long sz0 = m_n;
    long aveSz0 = (((sz0 - 1) * sz0) / 2) / binT;
    int step = 0;
    while (nextInternal < nn) {
      QDij minQ = qdA[0];
      if (!predictQ) minQ.q = DBL_MAX;
      for (int b = 0; b < numBins; ++b) binRange[b] = maxT2[b] = -DBL_MAX;
      for (int k = firstActive; k < nextInternal; k = next[k]) {
        int rk = redirect[k], bi = binLoc[rk];
        if (T[rk] > maxT2[bi]) {
          if (T[rk] > binRange[bi]) {
            maxT2[bi] = binRange[bi];
            binRange[bi] = T[rk];
          } else
            maxT2[bi] = T[rk];
        }
      }
      // candidate list (:1019-1068)
      int last = (int)adijA.size() - 1, ic;
      for (ic = last; ic >= 0; --ic) {
        ActiveDij& a = adijA[ic];
        if (a.active) break;
        if (redirect[a.i] != -1 && redirect[a.j] != -1) break;
        a.active = false;
      }
      if (ic != last) adijA.erase(adijA.begin() + (ic + 1), adijA.end());
      unsigned inactive = 0;
      for (ic = (int)adijA.size() - 1; ic >= 0; --ic) {
        ActiveDij& a = adijA[ic];
        if (!a.active) {
          inactive++;
          continue;
        }
        int ri = redirect[a.i], rj = redirect[a.j];
        if (ri == -1 || rj == -1)
          a.active = false;
        else {
          double q = ((double)a.d) * (m_n - 2) - T[ri] - T[rj];
          if (q <= minQ.q) {
            minQ.assign(q, a.d, a.i, a.j);
            qdA[n++].assign(q, a.d, a.i, a.j);
            if (n == QDASZ) {
              std::nth_element(qdA.begin(), qdA.begin() + QDA, qdA.begin() + n, QCmp());
              n = QDA;
            }
          }
        }
      }
      // candidate re-binning every candidateScan steps (:1071-1093)
      if ((n0 - m_n) % candidateScan == 0 && stepsUntilRebuild > candidateScan / 2) {
        for (ic = (int)adijA.size() - 1; ic >= 0; --ic) {
          ActiveDij& a = adijA[ic];
          if (!a.active) continue;
          int ri = redirect[a.i], rj = redirect[a.j];
          int bi = binLoc[ri], bj = binLoc[rj];
          double maxTSum = binRange[bi] + ((bi == bj) ? maxT2[bi] : binRange[bj]);
          double qLimit = ((double)a.d) * (m_n - 2) - maxTSum;
          if (qLimit > minQ.q) {
            a.active = false;
            pqA[idxA[bi][bj]].push(Dij{a.d, a.i, a.j});
          }
        }
      }
      if (adijA.size() > 0 && inactive > adijA.size() / 5) {
        auto goodEnd = std::partition(adijA.begin(), adijA.end(),
                                      [](const ActiveDij& a) { return a.active; });
        adijA.erase(goodEnd, adijA.end());
      }
      // bins (:1102-1144)
      for (int jb = 0; jb < binT; ++jb) {
        int lbin = binOrder[jb];
        PQ& pq = pqA[lbin];
        int bi = binPair[lbin].first, bj = binPair[lbin].second;
        double maxTSum = binRange[bi] + ((bi == bj) ? maxT2[bi] : binRange[bj]);
        while (!pq.empty()) {
          Dij dij = pq.top();
          int ri = redirect[dij.i], rj = redirect[dij.j];
          if (ri == -1 || rj == -1) {
            pq.pop();
            continue;
          }
          double q = ((double)dij.d) * (m_n - 2);
          double qLimit = q - maxTSum;
          q -= T[ri] + T[rj];
          if (qLimit > minQ.q) break;
          qdA[n++].assign(q, dij.d, dij.i, dij.j);
          if (n == QDASZ) {
            std::nth_element(qdA.begin(), qdA.begin() + QDA, qdA.begin() + n, QCmp());
            n = QDA;
          }
          if (q <= minQ.q) minQ.assign(q, dij.d, dij.i, dij.j);
          adijA.push_back(ActiveDij{dij.d, dij.i, dij.j, true});
          pq.pop();
        }
      }
      std::nth_element(qdA.begin(), qdA.begin() + std::min(n, QDA), qdA.begin() + n, QCmp());
      for (int jb = n; jb < QDA; ++jb) qdA[jb].q = DBL_MAX;
      for (int k = 0; k < QDA; ++k) qdAg[k] = qdA[k];  // allgather of one rank
      minQ.q = DBL_MAX;
      int minI = INT_MAX, minJ = INT_MAX;
      for (int ja = 0; ja < QDA; ++ja) {
        if (qdAg[ja].q <= minQ.q) {
          int newI = std::min(qdAg[ja].i, qdAg[ja].j);
          int newJ = std::max(qdAg[ja].i, qdAg[ja].j);
          if ((qdAg[ja].q < minQ.q) || (minJ < newJ || ((minJ == newJ) && (minI < newI)))) {
            minQ = qdAg[ja];
            minI = std::min(minQ.i, minQ.j);
            minJ = std::max(minQ.i, minQ.j);
          }
        }
      }
      int ii = std::min(minQ.i, minQ.j);
      minQ.j = std::max(minQ.i, minQ.j);
      minQ.i = ii;
      int riMin = redirect[minQ.i];
      {
        orc_join_t& J = joins[step++];
        J.node = nextInternal;
        J.left = minQ.i;
        J.right = minQ.j;
        J.d = minQ.d;
        J.Ti = T[redirect[minQ.i]];
        J.Tj = T[redirect[minQ.j]];
        J.n = m_n;
        J.q = minQ.q;
        J.lenL = branch(minQ.d, J.Ti, J.Tj, m_n, &J.lenR);
      }
      modifyMatrix(minQ, riMin);
      // PQ-size trigger (:1207-1223)
      if (--pq_ratio_ck == 0) {
        pq_ratio_ck = pqRatioInit;
        if (stepsUntilRebuild > 0) {
          long totalPQ = 0;
          for (auto& q : pqA) totalPQ += (long)q.size();
          totalPQ = totalPQ / (long)binT;
          double r = ((double)totalPQ) / ((double)aveSz0);
          if (r > pqRatio) {
            stepsUntilRebuild = 0;
            pqTriggers++;
          }
        }
      }
      m_n--;
      if (stepsUntilRebuild == 0) {
        redirect[nextInternal] = riMin;
        nextInternal++;
        adijA.clear();
        buildRowSums();
        buildBins();
        rebuilds++;
        stepsUntilRebuild = (m_n < minRebuild) ? INT_MAX : (int)(m_n * ratio);
        sz0 = m_n;
        aveSz0 = (((sz0 - 1) * sz0) / 2) / binT;
      } else {
        stepsUntilRebuild--;
        for (int b = 0; b < numBins; ++b)
          if (T[riMin] <= binRange[b]) {
            binLoc[riMin] = b;
            break;
          }
        int bi = binLoc[riMin];
        redirect[nextInternal] = riMin;
        for (int k = firstActive; k < nextInternal; k = next[k]) {
          int rk = redirect[k], bj = binLoc[rk];
          int a, b;
          if (rk < riMin) {
            a = k;
            b = nextInternal;
          } else {
            a = nextInternal;
            b = k;
          }
          pqA[idxA[bi][bj]].push(Dij{vd[rk], a, b});
        }
        nextInternal++;
      }
      // prediction for the next step (:1301-1322)
      for (QDij& qd : qdAg) {
        int ri, rj;
        if (qd.i < 0 || qd.j < 0 || (ri = redirect[qd.i]) == -1 || (rj = redirect[qd.j]) == -1) {
          qd.q = DBL_MAX;
          continue;
        }
        qd.q = ((double)qd.d) * (m_n - 2) - T[ri] - T[rj];
      }
      int szQ = std::min(2 * QDA, (int)qdAg.size());
      std::partial_sort(qdAg.begin(), qdAg.begin() + szQ, qdAg.end(), QCmp());
      qdA[0] = qdAg[0];
      int jc = 0;
      for (int jb = 1; jb < qdAg_sz; ++jb) {
        if (((qdA[jc].i != qdAg[jb].i) || (qdA[jc].j != qdAg[jb].j)) && (qdAg[jb].q != DBL_MAX)) {
          qdA[++jc] = qdAg[jb];
          if (jc == QDA - 1) break;
        }
      }
      n = jc + 1;
    }
  }
};

}  // namespace

extern "C" int orc_nj_reference(const float* D, int n, int numBins, float ratio,
                                float pqRatio, int qda, int predictQ,
                                orc_join_t* joins, int* pq_trigger) {
  NJ nj;
  nj.numBins = numBins;
  nj.ratio = ratio;
  nj.pqRatio = pqRatio;
  nj.QDA = qda;
  nj.predictQ = predictQ != 0;
  nj.init(D, n);
  if (n >= 2) nj.run(joins);
  if (pq_trigger) *pq_trigger = nj.pqTriggers;
  return nj.rebuilds;
}

// Textbook NJ with the same arithmetic: the whole upper triangle of q is scanned
// each step with q = (d*(n-2) - T_a) - T_b (operands in the order the
// reference's candidate list holds the pair) and the tie rule of
// SpeciesBlock.cpp:1169-1178; T follows the reference's incremental update and
// rebuild schedule (:864,1225-1247) without the heap-size trigger.
extern "C" int orc_nj_fullscan(const float* Din, int n0, float ratio,
                               orc_join_t* joins) {
  const int minRebuild = 200;
  std::vector<std::vector<float>> D(n0, std::vector<float>(n0));
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n0; ++j) D[i][j] = Din[(size_t)i * n0 + j];
  std::vector<int> active(n0), store(2 * n0, -1);  // ids in increasing order
  for (int i = 0; i < n0; ++i) active[i] = store[i] = i;
  std::vector<double> T(n0, 0.0);
  auto rowsums = [&]() {
    for (int a : active) {
      double s = 0.0;
      for (int b : active) s += D[store[a]][store[b]];
      T[store[a]] = s;
    }
  };
  rowsums();
  int steps = std::max((int)(n0 * ratio), minRebuild);
  int m = n0, rebuilds = 0, R = n0;
  for (int step = 0; step < n0 - 1; ++step) {
    double best = DBL_MAX;
    int bi = -1, bj = -1;
    for (size_t x = 0; x < active.size(); ++x)
      for (size_t y = x + 1; y < active.size(); ++y) {
        int I = active[x], Jd = active[y];
        int ri = store[I], rj = store[Jd];
        // operand order of the reference's candidate list: id order for pairs
        // made by the last bin rebuild, storage order for pairs made by the
        // insertion of a node created after it (SpeciesBlock.cpp:1280-1287)
        double dq = ((double)D[ri][rj]) * (m - 2);
        double q = (Jd >= R && rj < ri) ? (dq - T[rj]) - T[ri] : (dq - T[ri]) - T[rj];
        if (q < best || (q == best && (Jd > bj || (Jd == bj && I > bi)))) {
          best = q;
          bi = I;
          bj = Jd;
        }
      }
    int ri = store[bi], rj = store[bj];
    orc_join_t& J = joins[step];
    J.node = n0 + step;
    J.left = bi;
    J.right = bj;
    J.d = D[ri][rj];
    J.Ti = T[ri];
    J.Tj = T[rj];
    J.n = m;
    J.q = best;
    J.lenL = NJ::branch(J.d, J.Ti, J.Tj, m, &J.lenR);
    std::vector<int> rest;
    for (int a : active)
      if (a != bi && a != bj) rest.push_back(a);
    double sum = 0.0;
    for (int a : rest) {
      int rk = store[a];
      float Dik = D[ri][rk], Djk = D[rj][rk];
      float tmp = ((Dik + Djk) - J.d) / 2;
      sum += tmp;
      T[rk] += (double)(float)(tmp - (Dik + Djk));
      D[ri][rk] = tmp;
      D[rk][ri] = tmp;
    }
    T[ri] = sum;
    store[J.node] = ri;
    store[bi] = store[bj] = -1;
    rest.push_back(J.node);
    active.swap(rest);
    --m;
    if (steps == 0) {
      rowsums();
      ++rebuilds;
      R = J.node + 1;
      steps = (m < minRebuild) ? INT_MAX : (int)(m * ratio);
    } else {
      --steps;
    }
  }
  return rebuilds;
}

namespace {
void tree_text_rec(int node, const std::vector<int>& L, const std::vector<int>& R,
                   const std::vector<double>& len, int n0,
                   const char* const* names, std::string& s) {
  std::string lens;
  if (len[node] != DBL_MAX) {
    char buf[64];
    // TreeNode.cpp:68: setprecision(4) << showpoint << setw(8) is %#8.4g
    snprintf(buf, sizeof buf, ":%#8.4g", len[node]);
    lens = buf;
  }
  if (node < n0) {
    s += names[node];
    s += lens;
  } else {
    s += "(\n";
    tree_text_rec(L[node], L, R, len, n0, names, s);
    s += ",\n";
    tree_text_rec(R[node], L, R, len, n0, names, s);
    s += ")\n";
    s += lens;
  }
}
}  // namespace

extern "C" int orc_tree_text(const orc_join_t* joins, int n0,
                             const char* const* names, char* out, int cap) {
  int nn = 2 * n0 - 1;
  std::vector<int> L(nn, -1), R(nn, -1);
  std::vector<double> len(nn, DBL_MAX);
  for (int k = 0; k < n0 - 1; ++k) {
    const orc_join_t& J = joins[k];
    L[J.node] = J.left;
    R[J.node] = J.right;
    len[J.left] = J.lenL;
    len[J.right] = J.lenR;
  }
  std::string s;
  tree_text_rec(nn - 1, L, R, len, n0, names, s);
  s += ";\n";
  if ((int)s.size() + 1 <= cap) memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size() + 1;
}
