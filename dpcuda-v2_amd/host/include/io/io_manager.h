// io/io_manager.h at the reference's include path: class IOManager (io.hpp).
#pragma once
#include "../host_api.hpp"
