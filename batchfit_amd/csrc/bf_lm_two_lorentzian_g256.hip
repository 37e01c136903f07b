// bf_lm_two_lorentzian_g256.hip — LM kernels for model id 3, a workgroup of 256 lanes per fit.
#define BF_MODEL 3
#define BF_GROUP 256
#include "bf_instantiate.hpp"
