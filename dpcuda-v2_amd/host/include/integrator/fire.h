// integrator/fire.h at the reference's include path: class Fire, get_fire_config_dict (integrator.hpp).
#pragma once
#include "../host_api.hpp"
