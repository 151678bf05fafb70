// io/state_log.h at the reference's include path: the state log group (io.hpp).
#pragma once
#include "../host_api.hpp"
