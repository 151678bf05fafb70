// routines/minimization.h at the reference's include path: minimizeFire / minimizeAdam (routines.hpp).
#pragma once
#include "../host_api.hpp"
