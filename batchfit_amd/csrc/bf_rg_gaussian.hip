// bf_rg_gaussian.hip — ragged-batch kernels for model id 0.
#define BF_MODEL 0
#include "bf_ragged_instantiate.hpp"
