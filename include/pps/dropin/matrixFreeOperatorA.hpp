// Drop-in for the reference's matrixFreeOperatorA.hpp (MatrixFreeOperatorA: the operator lives in the GPU library), see pps_dropin.hpp.
#pragma once
#include "pps_dropin.hpp"
