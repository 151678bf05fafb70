// data/array_data.h at the reference's include path: struct ArrayData (array_data.hpp).
#pragma once
#include "../host_api.hpp"
