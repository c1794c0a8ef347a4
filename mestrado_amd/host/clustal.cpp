// Entry point with the phase sequence of the reference's wnj.cpp (:16-57) on
// one process and one device: options, distance matrix, guide tree,
// progressive alignment, timers.
#include <cstdio>

#include "CmdLineOptions.h"
#include "Runtime.h"
#include "SpeciesBlock.h"

int main(int argc, char* argv[]) {
  CmdLineOptions cmd(argc, argv);
  CmdLineOptions::state_t state = cmd.state();
  if (state != CmdLineOptions::iGOOD) return (state == CmdLineOptions::iBAD);

  SpeciesBlock sb(cmd);
  printf("readFasta %d\n", 0);
  sb.readFasta();
  printf("buildTree %d\n", 0);
  sb.buildTree();
  printf("freeMemory %d\n", 0);
  sb.freeMemory();
  printf("buildAlign %d\n", 0);
  sb.buildAlign();
  sb.reportTimers();
  mstrt::shutdown();
  return 0;
}
