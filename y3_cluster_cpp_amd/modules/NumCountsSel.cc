// NumCountsSel module library: the CosmoSIS entry points setup / execute /
// cleanup for the NumCountsSel evaluator (evaluators.hpp).
#include "evaluators.hpp"

DEFINE_COSMOSIS_HIP_EVALUATOR_MODULE(NumCountsSel)
