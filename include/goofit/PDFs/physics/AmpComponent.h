#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

/// A component of an amplitude model; evaluated only inside its parent.
class AmpComponent : public GooPdf {
  public:
    using GooPdf::GooPdf;
};

}  // namespace GooFit
