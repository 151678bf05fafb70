// io/arg_parsing.h at the reference's include path: load_configs (io.hpp).
#pragma once
#include "../host_api.hpp"
