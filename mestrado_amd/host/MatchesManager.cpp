#include "MatchesManager.h"

MatchesManager::MatchesManager(int sequencesCount, int wSize, int maxLen)
    : windowSize(wSize > 0 ? wSize : 1), windowsXYCount(sequencesCount), maxSequenceLength(maxLen) {}

MatchesManager::~MatchesManager() {}

void MatchesManager::put(int x, int y, const std::string& gx, const std::string& gy) {
  std::string v = gx;
  v.push_back('\0');
  v += gy;
  v.push_back('\0');
  if (x < y) {  // keyed as (max, min) like getSequence
    std::string w = gy;
    w.push_back('\0');
    w += gx;
    w.push_back('\0');
    store[std::make_pair(y, x)] = w;
  } else {
    store[std::make_pair(x, y)] = v;
  }
}

char* MatchesManager::getSequence(int x, int y) {
  if (x < y) return getSequence(y, x);
  auto it = store.find(std::make_pair(x, y));
  return it == store.end() ? nullptr : &it->second[0];
}

char* MatchesManager::getWindow(int windowX, int windowY) {
  const int x0 = windowX * windowSize, y0 = windowY * windowSize;
  for (auto& kv : store) {
    const int x = kv.first.first, y = kv.first.second;
    if (x >= x0 && x < x0 + windowSize && y >= y0 && y < y0 + windowSize) return &kv.second[0];
  }
  return nullptr;
}
