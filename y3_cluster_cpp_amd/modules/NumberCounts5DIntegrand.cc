// NumberCounts5DIntegrand module library: the CosmoSIS entry points setup / execute /
// cleanup for the NumberCounts5DIntegrand integrand (integrands.hpp).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(NumberCounts5DIntegrand)
