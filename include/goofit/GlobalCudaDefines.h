// Basic numeric definitions of the GooFit API (include/goofit/GlobalCudaDefines.h of
// the reference): FP64 throughout.
#pragma once

#include <cmath>

namespace GooFit {
using fptype = double;
}  // namespace GooFit

#define root2 1.4142135623730951
#define invRootPi 0.5641895835477563
