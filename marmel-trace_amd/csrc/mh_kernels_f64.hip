// fp64 instantiation (parity path). Built with -ffp-contract=off so every
// double operation rounds as in the reference C code.
#include "mh_kernels_impl.h"
using namespace mh;
MH_INSTANTIATE(double)
