// Drop-in for the reference's blockGrid.hpp (BlockGrid), see pps_dropin.hpp.
#pragma once
#include "pps_dropin.hpp"
