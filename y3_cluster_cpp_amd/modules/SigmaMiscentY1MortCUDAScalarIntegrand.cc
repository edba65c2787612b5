// SigmaMiscentY1MortCUDAScalarIntegrand module library: the CosmoSIS entry
// points setup / execute / cleanup for the real-data Delta-Sigma integrand
// (integrands.hpp).
#include "integrands.hpp"

DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(SigmaMiscentY1MortCUDAScalarIntegrand)
