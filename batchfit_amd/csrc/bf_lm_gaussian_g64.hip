// bf_lm_gaussian_g64.hip — LM kernels for model id 0, 64 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 64
#include "bf_instantiate.hpp"
