// bf_lm_gaussian_g256.hip — LM kernels for model id 0, a workgroup of 256 lanes per fit.
#define BF_MODEL 0
#define BF_GROUP 256
#include "bf_instantiate.hpp"
