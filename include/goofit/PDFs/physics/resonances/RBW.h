// Relativistic Breit-Wigner with Blatt-Weisskopf factors (resonances/RBW.h of the reference).
#pragma once

#include <goofit/PDFs/physics/resonances/Resonance.h>

namespace GooFit {
namespace Resonances {

class RBW : public ResonancePdf {
  public:
    RBW(std::string name, Variable ar, Variable ai, Variable mass, Variable width, unsigned int sp, unsigned int cyc,
        bool norm = true, bool sym = false);
    ~RBW() override = default;
};

}  // namespace Resonances
}  // namespace GooFit
