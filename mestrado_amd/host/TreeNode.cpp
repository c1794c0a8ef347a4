#include "TreeNode.h"

#include <cfloat>
#include <cstdio>

TreeNode::TreeNode()
    : index(-1), m_aligned(NALIGNED), m_length(DBL_MAX), mp_left(nullptr), mp_right(nullptr) {}

TreeNode::~TreeNode() {}

void TreeNode::init(int iNode, const char* nm) {
  m_name = nm;
  index = iNode;
  m_aligned = ALIGNED;
}

void TreeNode::setLength(double x) { m_length = x; }

void TreeNode::join(int iNode, TreeNode* left, TreeNode* right, float minD, double Ti, double Tj,
                    int nSpecies) {
  mp_left = left;
  mp_right = right;
  const double half = 0.5;
  double l, r;
  if (nSpecies == 2) {
    l = r = minD * half;
  } else {
    l = half * (minD + (Ti - Tj) / (nSpecies - 2));
    r = half * (minD + (Tj - Ti) / (nSpecies - 2));
  }
  // a negative side becomes zero and the other side takes up the difference
  if (l < 0.0) {
    r -= l;
    l = 0.0;
  } else if (r < 0.0) {
    l -= r;
    r = 0.0;
  }
  index = iNode;
  mp_left->setLength(l);
  mp_right->setLength(r);
}

void TreeNode::buildTreeOutput(std::ostringstream& s) {
  // Length suffix ":%#8.4g" (iostream: setprecision(4), showpoint, setw(8));
  // the root carries none. Written iteratively so deep trees cannot exhaust
  // the stack.
  struct Frame {
    TreeNode* n;
    int stage;
  };
  std::vector<Frame> st;
  st.push_back({this, 0});
  auto len = [](const TreeNode* n) {
    if (n->m_length == DBL_MAX) return std::string();
    char buf[64];
    snprintf(buf, sizeof buf, ":%#8.4g", n->m_length);
    return std::string(buf);
  };
  while (!st.empty()) {
    Frame& f = st.back();
    TreeNode* n = f.n;
    if (n->mp_left == nullptr) {
      s << n->m_name << len(n);
      st.pop_back();
      continue;
    }
    if (f.stage == 0) {
      s << "(\n";
      f.stage = 1;
      st.push_back({n->mp_left, 0});
    } else if (f.stage == 1) {
      s << ",\n";
      f.stage = 2;
      st.push_back({n->mp_right, 0});
    } else {
      s << ")\n" << len(n);
      st.pop_back();
    }
  }
}

void TreeNode::update() {
  if (m_aligned > NALIGNED) m_aligned = ALIGNED;
}

bool TreeNode::getNodes(std::vector<int>& nodes, int iNodes) {
  // Post-order collection of (parent, left, right) merges whose children are
  // both aligned, at most iNodes-1 of them (TreeNode.cpp:90-117).
  if ((mp_left->m_aligned == NALIGNED) && mp_left->getNodes(nodes, iNodes)) return true;
  if ((mp_right->m_aligned == NALIGNED) && mp_right->getNodes(nodes, iNodes)) return true;
  if ((mp_left->m_aligned > CUTTED) && (mp_right->m_aligned > CUTTED)) {
    if ((int)(nodes.size() / 3) < (iNodes - 1)) {
      nodes.push_back(index);
      nodes.push_back(mp_left->index);
      nodes.push_back(mp_right->index);
      m_aligned = ALIGNING;
    } else {
      mp_left->m_aligned = CUTTED;
      return true;
    }
  }
  return false;
}
