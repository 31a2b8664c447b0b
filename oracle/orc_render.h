// CPU restatement of the reference's shading and integrators.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include "orc_scene.h"
