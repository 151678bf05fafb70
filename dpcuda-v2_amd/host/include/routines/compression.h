// routines/compression.h at the reference's include path: jam_adam, compress_adam, decompress_adam, compress_to_phi_adam (routines.hpp).
#pragma once
#include "../host_api.hpp"
