// particles/rigid_bumpy/rigid_bumpy.h at the reference's include path: class RigidBumpy (particle.hpp).
#pragma once
#include "../../host_api.hpp"
