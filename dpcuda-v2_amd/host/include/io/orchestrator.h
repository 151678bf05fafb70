// io/orchestrator.h at the reference's include path: class Orchestrator (io.hpp).
#pragma once
#include "../host_api.hpp"
