// integrator/adam.h at the reference's include path: class Adam, get_adam_config_dict (integrator.hpp).
#pragma once
#include "../host_api.hpp"
