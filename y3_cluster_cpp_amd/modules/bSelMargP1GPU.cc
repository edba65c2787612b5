// bSelMargP1GPU module library: the CosmoSIS entry points setup / execute /
// cleanup for the Costanzi-2026 P[X] operator integrand with the P1 weight
// (integrands.hpp; p_operator_gpu_t.cuh:327-366).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(bSelMargP1GPU)
