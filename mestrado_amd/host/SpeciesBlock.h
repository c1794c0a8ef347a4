// Pipeline driver with the interface of the reference's SpeciesBlock
// (SpeciesBlock.h:37-58) for one process and one device: readFasta builds the
// score/"distance" matrix (SpeciesBlock.cpp:230-411), buildTree runs neighbour
// joining and writes <basename>.tree (:844-1343), buildAlign runs the
// progressive alignment over every merge of the guide tree (:1379-1591),
// reportTimers prints "BuildMtx,BuildTree,BuildAlign,Total" seconds to stderr
// (:104-116), measured as wall time. With -i m (CmdLineOptions.cpp:41,88-98
// accepts it; the reference reads nothing) the matrix a -o m run wrote is read
// instead of a FASTA file, and only the tree is built.
#ifndef MST_SPECIESBLOCK_H
#define MST_SPECIESBLOCK_H

#include <chrono>
#include <string>
#include <vector>

#include "CmdLineOptions.h"
#include "TreeNode.h"
#include "mestrado.h"

class SpeciesBlock {
 public:
  enum { tBuildMtx = 0, tBuildTree, tBuildAlign, tTotal };

  explicit SpeciesBlock(const CmdLineOptions& cmd);
  ~SpeciesBlock();

  void readFasta();
  void buildTree();
  void buildAlign();
  void freeMemory();
  void reportTimers();

  // build accessors
  const std::vector<mst_join>& joins() const { return m_joins; }
  const std::vector<std::string>& sequences() const { return seqs; }

 private:
  struct Timer {
    double acc = 0.0;
    std::chrono::steady_clock::time_point t0;
    void start() { t0 = std::chrono::steady_clock::now(); }
    void end() {
      acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  };

  CmdLineOptions m_cmd;
  void readMatrix();
  int m_n0;
  int maxSeqLen;
  std::vector<int> seqsLen;
  std::vector<std::string> seqsName;
  std::vector<std::string> seqs;
  std::vector<std::vector<float>> m_distances;
  std::vector<mst_join> m_joins;
  std::vector<TreeNode> m_nodes;
  Timer totalTime;
  Timer m_Timer[3];
};

#endif
