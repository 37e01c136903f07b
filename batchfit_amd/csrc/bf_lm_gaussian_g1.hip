// bf_lm_gaussian_g1.hip — LM kernels for model id 0, 1 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 1
#include "bf_instantiate.hpp"
