// host_api.hpp -- the whole C++ host API of this library in one header, for the headers
// of this directory, which sit at the include paths of the reference
// (include/particles/..., include/integrator/..., include/io/..., include/routines/...,
// include/utils/..., include/data/...) so that a driver written against the reference's
// headers compiles unchanged with -I <this directory>.  It also brings in the standard
// headers the reference's headers make visible to such drivers.
#pragma once
#include <chrono>
#include <cmath>
#include <filesystem>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../array_data.hpp"
#include "../config_dict.hpp"
#include "../factory.hpp"
#include "../integrator.hpp"
#include "../io.hpp"
#include "../particle.hpp"
#include "../routines.hpp"
#include "../standard_configs.hpp"
