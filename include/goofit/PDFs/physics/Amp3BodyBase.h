#pragma once

#include <goofit/PDFs/CombinePdf.h>

namespace GooFit {

class Amp3BodyBase : public CombinePdf {
  public:
    using CombinePdf::CombinePdf;
};

}  // namespace GooFit
