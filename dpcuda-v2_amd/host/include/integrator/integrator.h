// integrator/integrator.h at the reference's include path: class Integrator (integrator.hpp).
#pragma once
#include "../host_api.hpp"
