// Wall distance (apps/utils/calc_walldist.h), needed only by the mixing
// length and LES models, which are not part of this build.
#ifndef NSEM_FIELD_CALC_WALLDIST_H
#define NSEM_FIELD_CALC_WALLDIST_H

#include "types.h"

namespace Mesh {
void calc_walldist(Int, Int = 1);
}

#endif
