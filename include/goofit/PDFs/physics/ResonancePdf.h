#pragma once

#include <goofit/PDFs/physics/resonances/GS.h>
#include <goofit/PDFs/physics/resonances/NonRes.h>
#include <goofit/PDFs/physics/resonances/RBW.h>
