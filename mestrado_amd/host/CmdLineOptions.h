// Command-line options with the flags, defaults and validation of the
// reference's CmdLineOptions (CmdLineOptions.h:1-40, CmdLineOptions.cpp:31-142).
// Extra flags of this build (no counterpart in the reference) use letters the
// reference's getopt string leaves free: -b FILE (substitution matrix in the
// blosum64.txt format), -G n / -E n (gap open / extension, default 10 / 2),
// -f plain|batched (all-pairs form), -M (write <basename>.msa, the progressive
// alignment as gapped FASTA), -g n (HIP device).
#ifndef MST_CMDLINEOPTIONS_H
#define MST_CMDLINEOPTIONS_H

#include <string>

class CmdLineOptions {
 public:
  CmdLineOptions(int argc, char* argv[]);
  enum state_t { iBAD = 1, iHELP, iGOOD };

  state_t state() { return m_state; }

 private:
  char cleanup(const char* s);

 public:
  bool predictQ;
  bool allreduce;
  bool dupRemoval;
  int debugFlg;
  int clustSize;
  int qda;
  float rebuildRatio;
  float pqRatio;
  std::string inputFormat;
  std::string outputFormat;
  std::string inputFile;
  std::string baseName;
  // build extensions
  std::string blosumFile;
  int gapOpen, gapExt;
  int form;  // 0 plain, 1 batched
  bool writeMsa;
  int device;

 private:
  state_t m_state;
};

#endif
