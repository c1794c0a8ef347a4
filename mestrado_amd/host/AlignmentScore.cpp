#include "AlignmentScore.h"

#include <cstring>

#include "Runtime.h"

AlignmentScore::AlignmentScore(int gapOpen, int gapExt, unsigned int wSize, int maxSeqLen)
    : ctx(mstrt::context(gapOpen, gapExt)), windowSize(wSize), maxSeqLength(maxSeqLen), xSize(0),
      xSet(-1) {}

AlignmentScore::~AlignmentScore() {
  if (xSet >= 0) mst_seqset_free(ctx, xSet);
}

void AlignmentScore::parse(std::string& seqs, std::vector<int>& starts, std::vector<int>& lengths,
                           char* input, int size) {
  seqs.clear();
  lengths.clear();
  starts.assign(1, 0);
  const char* p = input + sizeof(int);  // sender rank
  for (int k = 0; k < size; ++k) {
    int len;
    p += sizeof(int);  // sequence number
    memcpy(&len, p, sizeof(int));
    p += sizeof(int);
    seqs.append(p, len);
    p += len + 1;  // residues and the terminating '\0'
    lengths.push_back(len);
    starts.push_back(starts.back() + len);
  }
}

static int upload(mst_ctx* ctx, const std::string& seqs, const std::vector<int>& starts) {
  std::vector<int64_t> off(starts.begin(), starts.end());
  int id = -1;
  mstrt::check(mst_seqset_upload(ctx, seqs.data(), off.data(), (int)starts.size() - 1, &id),
               "mst_seqset_upload");
  return id;
}

void AlignmentScore::loadDB(char* lbuf, int lSz) {
  xSize = lSz;
  parse(xSeqs, xStarts, xLengths, lbuf, xSize);
  if (xSet >= 0) mst_seqset_free(ctx, xSet);
  xSet = upload(ctx, xSeqs, xStarts);
}

void AlignmentScore::computeDistanceMtx(char* rbuf, int rSz, std::vector<std::vector<float>>& distances,
                                        int r0) {
  std::string ySeqs;
  std::vector<int> yStarts, yLengths;
  parse(ySeqs, yStarts, yLengths, rbuf, rSz);
  if (ySeqs == xSeqs && yStarts == xStarts) {
    // the local block against itself: every pair once, mirrored
    std::vector<int32_t> sc((size_t)xSize * xSize);
    mstrt::check(mst_all_pairs(ctx, xSet, mstrt::form(), sc.data(), nullptr), "mst_all_pairs");
    for (int x = 0; x < xSize; ++x)
      for (int y = 0; y < xSize; ++y) distances[x][r0 + y] = (float)sc[(size_t)x * xSize + y];
    return;
  }
  // a remote block: one set holding both, the cross pairs only
  std::string both = xSeqs + ySeqs;
  std::vector<int> st = xStarts;
  for (size_t k = 1; k < yStarts.size(); ++k) st.push_back(xStarts.back() + yStarts[k]);
  const int id = upload(ctx, both, st);
  std::vector<int32_t> pi, pj;
  for (int x = 0; x < xSize; ++x)
    for (int y = 0; y < rSz; ++y) {
      pi.push_back(x);
      pj.push_back(xSize + y);
    }
  std::vector<int32_t> out(pi.size());
  mstrt::check(mst_score_pairs(ctx, id, pi.data(), pj.data(), (int64_t)pi.size(), mstrt::form(),
                               out.data()),
               "mst_score_pairs");
  for (size_t k = 0; k < pi.size(); ++k) distances[pi[k]][r0 + pj[k] - xSize] = (float)out[k];
  mst_seqset_free(ctx, id);
}
