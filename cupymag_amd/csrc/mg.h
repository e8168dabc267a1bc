// Multigrid-preconditioned CG (see mg.hip).
#pragma once
#include <cstdint>

struct cpm_mg;
struct cpm_mat;

namespace cpm {

// PCG on the hierarchy's fine operator, on its context's stream.
int mg_pcg(cpm_mg* M, const double* b, double* x, int use_x0, double tol, int maxiter, int hint, int32_t* iters,
           double* resid, int* failed);
// the fine operator a hierarchy was built on
const cpm_mat* mg_operator(const cpm_mg* M);

}  // namespace cpm
