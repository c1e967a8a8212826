// Version check of a C++ caller linked against libtopotoolbox_amd.so, as
// the reference's test/versioninfo.cpp checks its own library: the header
// is the reference's topotoolbox.h when USE_REFERENCE_HEADER is defined
// (a caller relinked with no source change), else topotoolbox_amd.h.
extern "C" {
#ifdef USE_REFERENCE_HEADER
#include "topotoolbox.h"
#else
#include "topotoolbox_amd.h"
#endif
}

#include <iostream>

int main() {
  if (has_topotoolbox() != 1) return 1;
  std::cout << "topotoolbox v" << TOPOTOOLBOX_VERSION_MAJOR << "." << TOPOTOOLBOX_VERSION_MINOR
            << "." << TOPOTOOLBOX_VERSION_PATCH << std::endl;
  std::cout << "openmp v" << TOPOTOOLBOX_OPENMP_VERSION << std::endl;
  return 0;
}
