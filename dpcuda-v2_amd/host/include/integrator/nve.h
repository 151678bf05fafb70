// integrator/nve.h at the reference's include path: class NVE, get_nve_config_dict (integrator.hpp).
#pragma once
#include "../host_api.hpp"
