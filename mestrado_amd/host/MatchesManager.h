// Store of pairwise alignment strings with the interface of the reference's
// MatchesManager (MatchesManager.h:1-24): getSequence(x, y) is symmetric in
// (x, y) and getWindow(wx, wy) addresses a window of windowSize x windowSize
// pairs. The reference allocates count^2 buffers of maxSequenceLength bytes up
// front (MatchesManager.cpp:10-16); this store keeps only the pairs that were
// put, each as "gappedX\0gappedY\0".
#ifndef MST_MATCHESMANAGER_H
#define MST_MATCHESMANAGER_H

#include <map>
#include <string>
#include <utility>

class MatchesManager {
 public:
  MatchesManager(int sequencesCount, int windowSize, int maxSequenceLength);
  ~MatchesManager();

  // Alignment of sequences x and y ("gappedX\0gappedY\0" as put for the
  // ordered pair with x >= y), or nullptr when none was stored.
  char* getSequence(int x, int y);
  // First stored alignment of the window (windowX, windowY), or nullptr.
  char* getWindow(int windowX, int windowY);

  // build extension: store the alignment of (x, y)
  void put(int x, int y, const std::string& gx, const std::string& gy);
  int count() const { return (int)store.size(); }

 private:
  int windowSize;
  int windowsXYCount;
  int maxSequenceLength;
  std::map<std::pair<int, int>, std::string> store;
};

#endif
