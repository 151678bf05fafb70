// particles/base/particle.h at the reference's include path: class Particle (particle.hpp).
#pragma once
#include "../../host_api.hpp"
