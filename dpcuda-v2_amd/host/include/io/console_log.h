// io/console_log.h at the reference's include path: the console log group (io.hpp).
#pragma once
#include "../host_api.hpp"
