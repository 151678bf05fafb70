// routines/initialization.h at the reference's include path: the routine declarations (routines.hpp).
#pragma once
#include "../host_api.hpp"
