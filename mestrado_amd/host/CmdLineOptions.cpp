#include "CmdLineOptions.h"

#include <unistd.h>

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iostream>

static void printUsage(const char* execName) {
  std::cerr << "\nUsage:\n  " << execName << " [options] input_file\n\n"
            << "Options:\n"
            << "  -h, -?   : Print Usage\n"
            << "  -i style : Input style:  'a' for amino sequence data in Fasta format.\n"
            << "\t\t\t     'd' for dna sequence data in Fasta.\n"
            << "\t\t\t     'j' for dna sequence data in Fasta (using\n"
            << "\t\t\t\t Jukes-Cantor conversion).\n"
            << "\t\t\t     'k' for dna sequence data in Fasta (using\n"
            << "\t\t\t\t Kimura2 conversion).\n"
            << "\t\t\t     'm' for distance matrix.\n"
            << "  -o style : Output style: 't' tree output (default).\n"
            << "\t\t\t     'm' for distance matrix.\n"
            << "\t\t\t     'n' output a sequence with no duplicates.\n"
            << "\n"
            << "Expert Options:\n"
            << "  -k\t     : keep duplicates, do not remove early\n"
            << "  -R\t     : Use MPI_Allreduce instead of MPI_Allgather\n"
            << "  -d\tn    : Debug Level (1, 2, 3)\n"
            << "  -s\tn    : Cluster Size  (default = 30)\n"
            << "  -r\tx    : Rebuilt Ratio (default = 0.5)\n"
            << "  -q\tn    : QDA Size (default = 8) (zero turns off q-prediction)\n"
            << "\n"
            << "MI355X build options:\n"
            << "  -b FILE  : substitution matrix (24 rows of 24 comma-separated ints,\n"
            << "             letter order ARNDCQEGHILKMFPSTWYVBZX*; default BLOSUM62)\n"
            << "  -G n     : gap open penalty (default 10)\n"
            << "  -E n     : gap extension penalty (default 2)\n"
            << "  -f form  : all-pairs form 'plain' or 'batched' (default batched)\n"
            << "  -M       : write the progressive alignment to <basename>.msa\n"
            << "  -g n     : HIP device (default 0)\n"
            << std::endl;
}

CmdLineOptions::CmdLineOptions(int argc, char* argv[]) : m_state(iBAD) {
  bool help = false;
  pqRatio = 2.0f;
  rebuildRatio = 0.5f;
  clustSize = 30;
  qda = 8;
  inputFormat = "a";
  outputFormat = "t";
  debugFlg = 0;
  predictQ = true;
  allreduce = false;
  dupRemoval = true;
  gapOpen = 10;
  gapExt = 2;
  form = 1;
  writeMsa = false;
  device = 0;

  // The reference's option string "Ri:kh?d:s:r:m:no:q:" plus this build's.
  int opt;
  while ((opt = getopt(argc, argv, "Ri:kh?d:s:r:m:no:q:b:G:E:f:Mg:")) != -1) {
    switch (opt) {
      case 'R': allreduce = true; break;
      case 'i': inputFormat = optarg; break;
      case 'o': outputFormat = cleanup(optarg); break;
      case 'k': dupRemoval = false; break;
      case '?':
      case 'h': help = true; break;
      case 's': clustSize = (int)strtol(optarg, nullptr, 10); break;
      case 'r': rebuildRatio = (float)strtod(optarg, nullptr); break;
      case 'm': pqRatio = (float)strtod(optarg, nullptr); break;
      case 'd': debugFlg = (int)strtol(optarg, nullptr, 10); break;
      case 'q': qda = (int)strtol(optarg, nullptr, 10); break;
      case 'b': blosumFile = optarg; break;
      case 'G': gapOpen = (int)strtol(optarg, nullptr, 10); break;
      case 'E': gapExt = (int)strtol(optarg, nullptr, 10); break;
      case 'f': form = (std::string(optarg) == "plain") ? 0 : 1; break;
      case 'M': writeMsa = true; break;
      case 'g': device = (int)strtol(optarg, nullptr, 10); break;
      default: break;
    }
  }

  if (inputFormat != "a" && inputFormat != "d" && inputFormat != "j" && inputFormat != "k" &&
      inputFormat != "m") {
    std::cerr << "Bad input format: " << inputFormat << std::endl;
    printUsage(argv[0]);
    return;
  }
  if (outputFormat != "n" && outputFormat != "t" && outputFormat != "m") {
    std::cerr << "Bad output format: " << outputFormat << std::endl;
    printUsage(argv[0]);
    return;
  }
  if (help || optind >= argc) {
    m_state = iHELP;
    if (!help && optind >= argc) std::cerr << "\nNo input file => Quitting\n\n";
    printUsage(argv[0]);
    return;
  }
  if (qda == 0) {
    predictQ = false;
    qda = 2;
  }
  inputFile = argv[optind];

  // base name: the file name without directory and extension
  const std::string::size_type end = inputFile.size();
  std::string::size_type k;
  std::string::size_type i = ((k = inputFile.find_last_of('/')) != std::string::npos) ? k : 0;
  std::string::size_type j = ((k = inputFile.find_last_of('.')) != std::string::npos) ? k : end;
  if (inputFile.substr(i, 1) == "/" && i + 1 < end) ++i;
  baseName = inputFile.substr(i, j - i);

  std::fstream f(inputFile.c_str(), std::fstream::in);
  if (!f.good()) {
    std::cerr << "\nUnable to open file: " << inputFile << "\n" << std::endl;
    printUsage(argv[0]);
    return;
  }
  f.close();
  m_state = iGOOD;
}

char CmdLineOptions::cleanup(const char* s) { return (char)tolower(s[0]); }
