// CPU restatement of the reference's shading and integrators.  TEST INFRASTRUCTURE ONLY.
#include "orc_render.h"
