// io/energy_log.h at the reference's include path: the energy log group (io.hpp).
#pragma once
#include "../host_api.hpp"
