// NumberCounts3DIntegrand module library: the CosmoSIS entry points setup / execute /
// cleanup for the NumberCounts3DIntegrand integrand (integrands.hpp).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(NumberCounts3DIntegrand)
