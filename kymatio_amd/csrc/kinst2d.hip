// One instantiation family of the 2D kernels, compiled once per (dtype,
// prime class) by build.py with -DKYM_T=<float|double> -DKYM_P=<7|17|31>
// -DKYM_PICK=<function name> so the families build in parallel.
#include "kernels2d.h"

Scat2DKernel<KYM_T> KYM_PICK(int kind, int threads) { return pick_scat2d_p<KYM_T, KYM_P>(kind, threads); }
