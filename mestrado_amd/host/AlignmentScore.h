// All-pairs score driver with the interface of the reference's AlignmentScore
// (AlignmentScore.cuh:10-38): the constructor takes the gap penalties, a
// window size (a tuning hint here) and the longest sequence; loadDB/parse take
// the sequence buffer format of SpeciesBlock.cpp:309-322
// ([int rank][{int seqNo, int len, char[len], '\0'}...]); computeDistanceMtx
// fills distances[x][r0 + y] = score(X = x of loadDB, Y = y of rbuf).
#ifndef MST_ALIGNMENTSCORE_H
#define MST_ALIGNMENTSCORE_H

#include <string>
#include <vector>

#include "mestrado.h"

class AlignmentScore {
 public:
  AlignmentScore(int gapOpen, int gapExt, unsigned int wSize, int maxSeqLen);
  ~AlignmentScore();

  void loadDB(char* lbuf, int lSz);
  // Decodes `size` sequences of the buffer into seqs/starts/lengths (ASCII,
  // forward order; the device encodes and reverses them).
  void parse(std::string& seqs, std::vector<int>& starts, std::vector<int>& lengths, char* input,
             int size);
  void computeDistanceMtx(char* rbuf, int rSz, std::vector<std::vector<float>>& distances, int r0);

  // build accessors
  mst_ctx* context() { return ctx; }
  int setId() const { return xSet; }  // resident set of the last loadDB

 private:
  mst_ctx* ctx;
  unsigned int windowSize;
  int maxSeqLength;
  int xSize;
  std::string xSeqs;
  std::vector<int> xStarts, xLengths;
  int xSet;
};

#endif
