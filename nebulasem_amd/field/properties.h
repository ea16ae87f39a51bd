// Fluid constants of the apps (apps/utils/properties.h), enrolled in the
// controls file's general block.
#ifndef NSEM_FIELD_PROPERTIES_H
#define NSEM_FIELD_PROPERTIES_H

#include "util.h"

namespace Fluid {
extern Scalar density;
extern Scalar viscosity;
extern Scalar Pr;
extern Scalar Prt;
extern Scalar beta;
extern Scalar T0;
extern Scalar P0;
extern Scalar cp;
extern Scalar cv;
void enroll(Util::ParamList& params);
}  // namespace Fluid

#endif
