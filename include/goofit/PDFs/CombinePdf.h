#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class CombinePdf : public GooPdf {
  public:
    using GooPdf::GooPdf;
};

}  // namespace GooFit
