// Solver set-up of the apps (apps/utils/wrapper.h): reads the controls
// file's general, refinement, decomposition and prepare blocks.
#ifndef NSEM_FIELD_WRAPPER_H
#define NSEM_FIELD_WRAPPER_H

#include <fstream>

namespace Solver {
extern std::ifstream input;
void Initialize(int argc, char* argv[]);
void Finalize();
}  // namespace Solver

#endif
