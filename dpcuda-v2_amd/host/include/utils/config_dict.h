// utils/config_dict.h at the reference's include path: class ConfigDict (config_dict.hpp).
#pragma once
#include "../host_api.hpp"
