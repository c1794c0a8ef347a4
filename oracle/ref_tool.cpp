// ORACLE — test infrastructure only. A small driver of our own around the
// reference's host classes TreeNode (TreeNode.cpp), FastaFile (FastaFile.cpp)
// and CmdLineOptions (CmdLineOptions.cpp), compiled from the reference's own
// sources by oracle/Makefile into oracle/_ref/ref_tool. Tests run it as a
// subprocess to compare the build's tree text, FASTA parsing and option
// parsing with the reference's.
//
//   ref_tool tree  < joins.txt     (first line "n0", then one name per line,
//                                   then n0-1 lines "node left right d Ti Tj n"
//                                   with d as %a float and Ti/Tj as %a double)
//   ref_tool fasta FILE            (prints "name\tlength\tsequence" per record)
//   ref_tool cmd ARGS...           (prints the parsed option fields)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "CmdLineOptions.h"
#include "FastaFile.h"
#include "TreeNode.h"

static int do_tree() {
  int n0 = 0;
  if (scanf("%d", &n0) != 1 || n0 < 1) return 2;
  std::vector<std::string> names(n0);
  char buf[4096];
  for (int i = 0; i < n0; ++i) {
    if (scanf(" %4095[^\n]", buf) != 1) return 2;
    names[i] = buf;
  }
  int nn = 2 * n0 - 1;
  std::vector<TreeNode> nodes(nn);
  for (int i = 0; i < n0; ++i) nodes[i].init(i, names[i].c_str());
  for (int k = 0; k < n0 - 1; ++k) {
    int node, l, r, n;
    char sd[64], sti[64], stj[64];
    if (scanf("%d %d %d %63s %63s %63s %d", &node, &l, &r, sd, sti, stj, &n) != 7) return 2;
    float d = strtof(sd, nullptr);
    double ti = strtod(sti, nullptr), tj = strtod(stj, nullptr);
    nodes[node].join(node, &nodes[l], &nodes[r], d, ti, tj, n);
  }
  std::ostringstream s;
  nodes[nn - 1].buildTreeOutput(s);
  std::cout << s.str() << ";\n";
  return 0;
}

static int do_fasta(const char* path) {
  FastaFile f(path);
  int len = 0;
  std::string seq = f.nextSeq(&len);
  while (len > 0) {
    std::cout << f.getSeqName() << "\t" << len << "\t" << seq << "\n";
    seq = f.nextSeq(&len);
  }
  return 0;
}

static int do_cmd(int argc, char** argv) {
  CmdLineOptions c(argc, argv);
  std::cout << "state=" << (int)c.state() << "\n";
  if (c.state() == CmdLineOptions::iGOOD) {
    std::cout << "predictQ=" << c.predictQ << "\nallreduce=" << c.allreduce
              << "\ndupRemoval=" << c.dupRemoval << "\ndebugFlg=" << c.debugFlg
              << "\nclustSize=" << c.clustSize << "\nqda=" << c.qda
              << "\nrebuildRatio=" << c.rebuildRatio << "\npqRatio=" << c.pqRatio
              << "\ninputFormat=" << c.inputFormat << "\noutputFormat=" << c.outputFormat
              << "\ninputFile=" << c.inputFile << "\nbaseName=" << c.baseName << "\n";
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (!strcmp(argv[1], "tree")) return do_tree();
  if (!strcmp(argv[1], "fasta") && argc == 3) return do_fasta(argv[2]);
  if (!strcmp(argv[1], "cmd")) return do_cmd(argc - 1, argv + 1);
  return 2;
}
