// Drop-in for the reference's communicationMPI.hpp (CommunicatorMPI: on one GPU the halo exchange is the device ghost fill), see pps_dropin.hpp.
#pragma once
#include "pps_dropin.hpp"
