// FASTA reader with the interface and behaviour of the reference's FastaFile
// (FastaFile.h:1-24, FastaFile.cpp:21-60): the name is the header line
// verbatim (with '>'), the sequence is the concatenation of the following
// lines, a header not starting with '>' or an empty sequence aborts, and a
// final line without a trailing newline is not part of the sequence.
#ifndef MST_FASTAFILE_H
#define MST_FASTAFILE_H

#include <fstream>
#include <string>

class FastaFile {
 public:
  explicit FastaFile(const char* file);
  ~FastaFile();

  std::string getSeqName();
  // Next record; *length = 0 at the end of the file.
  std::string nextSeq(int* length);

 private:
  int lineNumber;
  std::string name;
  std::fstream fstr;
};

#endif
