// Pairwise/profile alignment driver with the interface of the reference's
// AlignmentMatch (AlignmentMatch.cuh:9-49). computeAlignment() runs every
// merge of aSeqs (iNode: index = the new node, iSpecies = its two children,
// node ids as in TreeNode; children before parents) as a profile-profile
// alignment on the device and leaves in seqs[i] the gapped row of sequence i
// in the alignment of the last merge's subtree. The reference's version
// discards its results and never runs profile merges (AlignmentMatch.cu:289-
// 297,428-430).
#ifndef MST_ALIGNMENTMATCH_H
#define MST_ALIGNMENTMATCH_H

#include <string>
#include <vector>

#include "MatchesManager.h"
#include "mestrado.h"

struct iNode {
  int index;
  int vecSize;
  bool aligned;
  std::vector<int> iSpecies;
};

class AlignmentMatch {
 public:
  AlignmentMatch(int gapOpen, int gapExt, unsigned int wSize, int maxSeqLen);
  ~AlignmentMatch();

  // Keeps the sequences for the merges (the device encodes and reverses them).
  void prepare(std::vector<std::string>& seqs, int myProc);
  void computeAlignment(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs, int myProc);
  // Leaf-leaf merges of aSeqs as pairwise alignments, stored in matches().
  void computePairwise(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs, int myProc);
  // Every merge of aSeqs as a progressive alignment.
  void computeMSA(std::vector<iNode>& aSeqs, std::vector<std::string>& seqs, int myProc);

  MatchesManager& matches() { return *mm; }
  int msaLength() const { return msaLen; }

 private:
  mst_ctx* ctx;
  unsigned int windowSize;
  int maxSeqLength;
  int nSize;
  int set;
  int msaLen;
  MatchesManager* mm;
};

#endif
