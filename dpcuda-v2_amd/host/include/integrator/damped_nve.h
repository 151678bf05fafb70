// integrator/damped_nve.h at the reference's include path: the integrator declarations (integrator.hpp); the damped NVE integrator itself is outside this library's scope.
#pragma once
#include "../host_api.hpp"
