// FCNBase -- see Minuit2.h
#pragma once
#include <Minuit2/Minuit2.h>
