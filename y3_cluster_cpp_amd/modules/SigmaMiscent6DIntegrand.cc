// SigmaMiscent6DIntegrand module library: the CosmoSIS entry points setup / execute /
// cleanup for the SigmaMiscent6DIntegrand integrand (integrands.hpp).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(SigmaMiscent6DIntegrand)
