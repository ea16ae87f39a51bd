// Solve of the field API (src/solvers/solve.h): applyImplicitBCs, then a
// diagonal system by Su / ap or a symmetric FV system by the library's
// DIAG-preconditioned conjugate gradient on the device, then
// applyExplicitBCs (solve.cpp:563-593).
#ifndef NSEM_FIELD_SOLVE_H
#define NSEM_FIELD_SOLVE_H

#include "field.h"

void Solve(const MeshMatrix<Scalar>&);
void Solve(const MeshMatrix<Vector>&);
void Solve(const MeshMatrix<STensor>&);
void Solve(const MeshMatrix<Tensor>&);

#endif
