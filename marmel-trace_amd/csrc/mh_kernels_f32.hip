// fp32 instantiation (throughput path, FMA contraction allowed).
#include "mh_kernels_impl.h"
using namespace mh;
MH_INSTANTIATE(float)
