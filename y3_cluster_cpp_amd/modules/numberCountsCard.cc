// numberCountsCard module library: the CosmoSIS entry points setup / execute /
// cleanup for the real-data integrand class of the same name (integrands.hpp).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(numberCountsCard)
