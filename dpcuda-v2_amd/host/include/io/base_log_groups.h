// io/base_log_groups.h at the reference's include path: the log groups and config_from_names* (io.hpp).
#pragma once
#include "../host_api.hpp"
