#include "FastaFile.h"

#include <cstdlib>
#include <iostream>

FastaFile::FastaFile(const char* file) : lineNumber(0) { fstr.open(file, std::fstream::in); }

FastaFile::~FastaFile() {
  if (fstr.is_open()) fstr.close();
}

std::string FastaFile::getSeqName() { return name; }

std::string FastaFile::nextSeq(int* length) {
  std::string header;
  std::getline(fstr, header);
  ++lineNumber;
  if (fstr.eof()) {  // nothing after the last record (or a header without '\n')
    *length = 0;
    return std::string();
  }
  if (header.empty() || header[0] != '>') {
    std::cerr << "Name string does not start with \">\" at line: " << lineNumber << std::endl;
    abort();
  }
  name = header;
  std::string seq;
  for (std::string text; !fstr.eof() && fstr.peek() != '>';) {
    std::getline(fstr, text);
    ++lineNumber;
    if (!fstr.good()) break;  // a last line lacking '\n' is dropped, as in the reference
    seq += text;
  }
  if (seq.empty()) {
    std::cerr << "Wrong format at line: " << lineNumber << std::endl;
    abort();
  }
  *length = (int)seq.size();
  return seq;
}
