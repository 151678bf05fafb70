// particles/standard_configs.h at the reference's include path: the standard particle configurations (standard_configs.hpp).
#pragma once
#include "../host_api.hpp"
