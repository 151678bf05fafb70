// particles/factory.h at the reference's include path: createParticle / loadParticle (factory.hpp).
#pragma once
#include "../host_api.hpp"
