#pragma once

#include <goofit/fitting/FitManagerMinuit2.h>

namespace GooFit {
using FitManager = GooFit::FitManagerMinuit2;
}  // namespace GooFit
