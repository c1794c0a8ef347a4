// Guide-tree node with the interface of the reference's TreeNode
// (TreeNode.h:1-41): join() computes the branch lengths of neighbour joining
// (TreeNode.cpp:30-58), buildTreeOutput() writes the reference's tree text
// (TreeNode.cpp:60-81), getNodes()/update() schedule merges (TreeNode.cpp:83-117).
#ifndef MST_TREENODE_H
#define MST_TREENODE_H

#include <sstream>
#include <string>
#include <vector>

#define NALIGNED 0
#define CUTTED 1
#define ALIGNING 2
#define ALIGNED 3

class TreeNode {
 public:
  TreeNode();
  ~TreeNode();

  void update();
  void init(int iNode, const char* nm);
  void buildTreeOutput(std::ostringstream& s);
  bool getNodes(std::vector<int>& nodes, int iNodes);
  void join(int iNode, TreeNode* left, TreeNode* right, float minD, double Ti, double Tj,
            int nSpecies);

  // build accessors (not in the reference's interface)
  int getIndex() const { return index; }
  double getLength() const { return m_length; }
  TreeNode* left() const { return mp_left; }
  TreeNode* right() const { return mp_right; }
  const std::string& name() const { return m_name; }

 private:
  void setLength(double x);

  int index;
  short m_aligned;
  double m_length;
  std::string m_name;
  TreeNode* mp_left;
  TreeNode* mp_right;
};

#endif
