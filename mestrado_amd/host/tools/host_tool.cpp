// Test driver of the drop-in host classes (no device needed):
//   host_tool fasta FILE     "name\tlength\tsequence" per record (FastaFile)
//   host_tool cmd ARGS...    the parsed option fields (CmdLineOptions)
//   host_tool tree < joins   tree text (TreeNode); input format: n0, n0 names,
//                            then n0-1 lines "node left right d Ti Tj n"
//   host_tool matrix FILE    the 576 values of a matrix file (-b), one row of
//                            24 per line, or exit status 1
//   host_tool matches W < ops  MatchesManager with window size W: lines
//                            "put x y gappedX gappedY", "seq x y", "win wx wy";
//                            seq/win print the two strings stored or "null"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "CmdLineOptions.h"
#include "FastaFile.h"
#include "MatchesManager.h"
#include "Runtime.h"
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
  std::vector<TreeNode> nodes(2 * n0 - 1);
  for (int i = 0; i < n0; ++i) nodes[i].init(i, names[i].c_str());
  for (int k = 0; k < n0 - 1; ++k) {
    int node, l, r, n;
    char sd[64], sti[64], stj[64];
    if (scanf("%d %d %d %63s %63s %63s %d", &node, &l, &r, sd, sti, stj, &n) != 7) return 2;
    nodes[node].join(node, &nodes[l], &nodes[r], strtof(sd, nullptr), strtod(sti, nullptr),
                     strtod(stj, nullptr), n);
  }
  std::ostringstream s;
  nodes[2 * n0 - 2].buildTreeOutput(s);
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

static int do_matrix(const char* path) {
  int8_t m[576];
  if (!mstrt::loadMatrix(path, m)) return 1;
  for (int r = 0; r < 24; ++r)
    for (int c = 0; c < 24; ++c) std::cout << (int)m[24 * r + c] << (c == 23 ? "\n" : " ");
  return 0;
}

static void print_pair(const char* p) {
  if (!p) {
    std::cout << "null\n";
    return;
  }
  std::cout << p << " " << p + strlen(p) + 1 << "\n";
}

static int do_matches(int window) {
  MatchesManager mm(1 << 20, window, 0);
  std::string op;
  while (std::cin >> op) {
    int a, b;
    if (!(std::cin >> a >> b)) return 2;
    if (op == "put") {
      std::string gx, gy;
      if (!(std::cin >> gx >> gy)) return 2;
      mm.put(a, b, gx, gy);
    } else if (op == "seq") {
      print_pair(mm.getSequence(a, b));
    } else if (op == "win") {
      print_pair(mm.getWindow(a, b));
    } else {
      return 2;
    }
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (!strcmp(argv[1], "tree")) return do_tree();
  if (!strcmp(argv[1], "fasta") && argc == 3) return do_fasta(argv[2]);
  if (!strcmp(argv[1], "cmd")) return do_cmd(argc - 1, argv + 1);
  if (!strcmp(argv[1], "matrix") && argc == 3) return do_matrix(argv[2]);
  if (!strcmp(argv[1], "matches") && argc == 3) return do_matches(atoi(argv[2]));
  return 2;
}
