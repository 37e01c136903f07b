// bf_lm_pseudo_voigt_g256.hip — LM kernels for model id 2, a workgroup of 256 lanes per fit.
#define BF_MODEL 2
#define BF_GROUP 256
#include "bf_instantiate.hpp"
