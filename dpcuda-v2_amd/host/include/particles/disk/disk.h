// particles/disk/disk.h at the reference's include path: class Disk (particle.hpp).
#pragma once
#include "../../host_api.hpp"
