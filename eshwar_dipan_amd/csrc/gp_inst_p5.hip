// gp_inst_p5.hip -- kernels for P = 5 phases; compiled once per real type
// (-DGP_REAL=double with -ffp-contract=off, -DGP_REAL=float with -ffp-contract=fast).
#include "gp_run.hpp"

#ifndef GP_REAL
#error GP_REAL must be double or float
#endif

GP_INSTANTIATE(5, 1, GP_REAL)
GP_INSTANTIATE(5, 2, GP_REAL)
