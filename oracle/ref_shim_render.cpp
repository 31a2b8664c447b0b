// C wrapper over the reference's integrators (see ref_shim.cpp).  TEST INFRASTRUCTURE ONLY.
