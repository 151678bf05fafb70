// io/io_utils.h at the reference's include path: the file helpers (io.hpp).
#pragma once
#include "../host_api.hpp"
