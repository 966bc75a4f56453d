// pps_nn.hip -- learned multilevel preconditioner (placeholder translation unit)
#include "pps_nn.hpp"
