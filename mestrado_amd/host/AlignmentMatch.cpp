#include "AlignmentMatch.h"

#include <algorithm>
#include <map>

#include "Runtime.h"

AlignmentMatch::AlignmentMatch(int gapOpen, int gapExt, unsigned int wSize, int maxSeqLen)
    : ctx(mstrt::context(gapOpen, gapExt)), windowSize(wSize), maxSeqLength(maxSeqLen), nSize(0),
      set(-1), msaLen(0), mm(nullptr) {}

AlignmentMatch::~AlignmentMatch() {
  if (set >= 0) mst_seqset_free(ctx, set);
  delete mm;
}

void AlignmentMatch::prepare(std::vector<std::string>& seqs, int /*myProc*/) {
  if (set >= 0) mst_seqset_free(ctx, set);
  std::string all;
  std::vector<int64_t> off(1, 0);
  for (auto& s : seqs) {
    all += s;
    off.push_back(off.back() + (int64_t)s.size());
  }
  nSize = (int)seqs.size();
  mstrt::check(mst_seqset_upload(ctx, all.data(), off.data(), nSize, &set), "mst_seqset_upload");
  delete mm;
  mm = new MatchesManager(nSize, (int)windowSize, 2 * maxSeqLength);
}

void AlignmentMatch::computeAlignment(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs,
                                      int myProc) {
  computePairwise(aSeqs, seqs, myProc);
  computeMSA(aSeqs, seqs, myProc);
}

void AlignmentMatch::computePairwise(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs,
                                     int /*myProc*/) {
  if (set < 0) prepare(seqs, 0);
  std::vector<int32_t> pi, pj;
  for (auto& a : aSeqs)
    if (a.iSpecies.size() == 2 && a.iSpecies[0] < nSize && a.iSpecies[1] < nSize) {
      pi.push_back(a.iSpecies[0]);
      pj.push_back(a.iSpecies[1]);
    }
  if (pi.empty()) return;
  int64_t stride = 1;
  for (size_t k = 0; k < pi.size(); ++k)
    stride = std::max<int64_t>(stride, (int64_t)seqs[pi[k]].size() + (int64_t)seqs[pj[k]].size() + 1);
  std::vector<char> gx(pi.size() * stride), gy(pi.size() * stride);
  std::vector<int32_t> sc(pi.size());
  mstrt::check(mst_align_pairs(ctx, set, pi.data(), pj.data(), (int64_t)pi.size(), sc.data(),
                               gx.data(), gy.data(), stride),
               "mst_align_pairs");
  for (size_t k = 0; k < pi.size(); ++k) mm->put(pi[k], pj[k], &gx[k * stride], &gy[k * stride]);
}

void AlignmentMatch::computeMSA(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs,
                                int /*myProc*/) {
  if (set < 0) prepare(seqs, 0);
  if (nSize < 2 || aSeqs.empty()) return;
  // Joins in creation order of the merge nodes; node ids follow TreeNode.
  std::vector<iNode> merges;
  for (auto& a : aSeqs)
    if (a.iSpecies.size() == 2) merges.push_back(a);
  std::sort(merges.begin(), merges.end(),
            [](const iNode& a, const iNode& b) { return a.index < b.index; });
  if ((int)merges.size() != nSize - 1) {
    fprintf(stderr, "computeMSA: %d merges for %d sequences; a whole guide tree is needed\n",
            (int)merges.size(), nSize);
    exit(1);
  }
  std::vector<mst_join> joins(merges.size());
  for (size_t k = 0; k < merges.size(); ++k) {
    joins[k] = mst_join();
    joins[k].node = merges[k].index;
    joins[k].left = std::min(merges[k].iSpecies[0], merges[k].iSpecies[1]);
    joins[k].right = std::max(merges[k].iSpecies[0], merges[k].iSpecies[1]);
    merges[k].aligned = true;
  }
  mstrt::check(mst_msa(ctx, set, joins.data(), &msaLen), "mst_msa");
  const int64_t stride = msaLen + 1;
  std::vector<char> rows((size_t)nSize * stride);
  mstrt::check(mst_msa_rows(ctx, rows.data(), stride), "mst_msa_rows");
  for (int i = 0; i < nSize; ++i) seqs[i].assign(&rows[(size_t)i * stride], msaLen);
  for (auto& a : aSeqs) a.aligned = true;
}
