#include "SpeciesBlock.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "AlignmentMatch.h"
#include "AlignmentScore.h"
#include "FastaFile.h"
#include "Runtime.h"

SpeciesBlock::SpeciesBlock(const CmdLineOptions& cmd) : m_cmd(cmd), m_n0(0), maxSeqLen(0) {
  totalTime.start();
  int8_t subst[576];
  bool have = false;
  if (!cmd.blosumFile.empty()) {
    if (!mstrt::loadMatrix(cmd.blosumFile, subst)) {
      std::cerr << "Unable to read substitution matrix: " << cmd.blosumFile << std::endl;
      exit(1);
    }
    have = true;
  }
  mstrt::configure(have ? subst : nullptr, cmd.device, cmd.form);
}

SpeciesBlock::~SpeciesBlock() {}

// -i m: the matrix a run with -o m wrote (first line n, then per row the
// sequence name and its n values); the tree is built from it, and there are no
// sequences to align.
void SpeciesBlock::readMatrix() {
  std::ifstream in(m_cmd.inputFile.c_str());
  std::string line;
  int n = -1;
  if (!std::getline(in, line) || sscanf(line.c_str(), "%d", &n) != 1 || n < 0) {
    std::cerr << "Bad matrix file: " << m_cmd.inputFile << std::endl;
    exit(1);
  }
  m_distances.assign(n, std::vector<float>(n, 0.0f));
  for (int i = 0; i < n; ++i) {
    if (!std::getline(in, line)) {
      std::cerr << "Bad matrix file: " << m_cmd.inputFile << " (row " << i << " missing)" << std::endl;
      exit(1);
    }
    // the n values are the last n fields; the name (which may hold spaces) precedes them
    size_t end = line.size();
    for (int j = n - 1; j >= 0; --j) {
      while (end > 0 && line[end - 1] == ' ') --end;
      size_t b = line.rfind(' ', end == 0 ? 0 : end - 1);
      b = (b == std::string::npos) ? 0 : b + 1;
      char* stop = nullptr;
      m_distances[i][j] = strtof(line.c_str() + b, &stop);
      if (b >= end || stop != line.c_str() + end) {
        std::cerr << "Bad matrix file: " << m_cmd.inputFile << " (row " << i << ")" << std::endl;
        exit(1);
      }
      end = b;
    }
    while (end > 0 && line[end - 1] == ' ') --end;
    seqsName.push_back(line.substr(0, end));
  }
  m_n0 = n;
  m_nodes.assign(m_n0 > 0 ? 2 * m_n0 - 1 : 0, TreeNode());
  for (int i = 0; i < m_n0; ++i) m_nodes[i].init(i, seqsName[i].c_str());
}

void SpeciesBlock::readFasta() {
  m_Timer[tBuildMtx].start();
  if (m_cmd.inputFormat == "m") {
    readMatrix();
    m_Timer[tBuildMtx].end();
    return;
  }
  FastaFile file(m_cmd.inputFile.c_str());
  int len = 0;
  std::string seq = file.nextSeq(&len);
  while (len > 0) {
    seqs.push_back(seq);
    seqsName.push_back(file.getSeqName());
    seqsLen.push_back(len);
    if (maxSeqLen < len) maxSeqLen = len;
    seq = file.nextSeq(&len);
  }
  m_n0 = (int)seqs.size();
  m_nodes.assign(m_n0 > 0 ? 2 * m_n0 - 1 : 0, TreeNode());
  for (int i = 0; i < m_n0; ++i) m_nodes[i].init(i, seqsName[i].c_str());

  // the sequence buffer of SpeciesBlock.cpp:309-322 for this (only) process
  std::vector<char> buf(sizeof(int));
  int rank = 0;
  memcpy(buf.data(), &rank, sizeof(int));
  for (int i = 0; i < m_n0; ++i) {
    size_t o = buf.size();
    buf.resize(o + 2 * sizeof(int) + seqsLen[i] + 1);
    memcpy(&buf[o], &i, sizeof(int));
    memcpy(&buf[o + sizeof(int)], &seqsLen[i], sizeof(int));
    memcpy(&buf[o + 2 * sizeof(int)], seqs[i].data(), seqsLen[i]);
    buf[o + 2 * sizeof(int) + seqsLen[i]] = '\0';
  }
  m_distances.assign(m_n0, std::vector<float>(m_n0, 0.0f));
  if (m_n0 > 0) {
    AlignmentScore align(m_cmd.gapOpen, m_cmd.gapExt, 16, maxSeqLen);
    align.loadDB(buf.data(), m_n0);
    align.computeDistanceMtx(buf.data(), m_n0, m_distances, 0);
  }
  m_Timer[tBuildMtx].end();

  if (m_cmd.outputFormat == "m") {  // the score matrix the tree is built from
    std::ofstream o((m_cmd.baseName + ".mtx").c_str());
    o << m_n0 << "\n";
    char v[32];
    for (int i = 0; i < m_n0; ++i) {
      o << seqsName[i];
      for (int j = 0; j < m_n0; ++j) {
        snprintf(v, sizeof v, " %.9g", m_distances[i][j]);  // every float round-trips
        o << v;
      }
      o << "\n";
    }
  }
}

void SpeciesBlock::buildTree() {
  m_Timer[tBuildTree].start();
  if (m_n0 > 1) {
    std::vector<float> D((size_t)m_n0 * m_n0);
    for (int i = 0; i < m_n0; ++i)
      memcpy(&D[(size_t)i * m_n0], m_distances[i].data(), sizeof(float) * m_n0);
    m_joins.assign(m_n0 - 1, mst_join());
    mst_ctx* ctx = mstrt::context(m_cmd.gapOpen, m_cmd.gapExt);
    mstrt::check(mst_nj(ctx, D.data(), m_n0, m_cmd.rebuildRatio, m_joins.data()), "mst_nj");
    for (const mst_join& j : m_joins)
      m_nodes[j.node].join(j.node, &m_nodes[j.left], &m_nodes[j.right], j.d, j.Ti, j.Tj, j.n);
  }
  m_Timer[tBuildTree].end();
  if (m_n0 > 0) {
    std::ostringstream s;
    m_nodes[2 * m_n0 - 2].buildTreeOutput(s);
    std::ofstream o((m_cmd.baseName + ".tree").c_str());
    o << s.str() << ";\n";
  }
}

void SpeciesBlock::freeMemory() {
  m_distances.clear();
  m_distances.shrink_to_fit();
}

void SpeciesBlock::buildAlign() {
  m_Timer[tBuildAlign].start();
  if (m_n0 > 1 && (int)seqs.size() == m_n0) {
    AlignmentMatch align(m_cmd.gapOpen, m_cmd.gapExt, 16, maxSeqLen);
    align.prepare(seqs, 0);
    std::vector<int> nodes;
    m_nodes[2 * m_n0 - 2].getNodes(nodes, m_n0);
    std::vector<iNode> aSeqs;
    for (size_t k = 0; k + 2 < nodes.size(); k += 3) {
      iNode a;
      a.index = nodes[k];
      a.aligned = false;
      a.iSpecies = {nodes[k + 1], nodes[k + 2]};
      a.vecSize = (nodes[k + 1] < m_n0) + (nodes[k + 2] < m_n0);
      aSeqs.push_back(a);
    }
    align.computeAlignment(aSeqs, seqs, 0);
    for (int i = m_n0; i < 2 * m_n0 - 2; ++i) m_nodes[i].update();
  }
  m_Timer[tBuildAlign].end();
  if (m_cmd.writeMsa && m_n0 > 0 && (int)seqs.size() == m_n0) {
    std::ofstream o((m_cmd.baseName + ".msa").c_str());
    for (int i = 0; i < m_n0; ++i) o << seqsName[i] << "\n" << seqs[i] << "\n";
  }
}

void SpeciesBlock::reportTimers() {
  totalTime.end();
  for (int k = 0; k < 3; ++k) fprintf(stderr, "%5.4f,", m_Timer[k].acc);
  fprintf(stderr, "%5.4f\n", totalTime.acc);
}
