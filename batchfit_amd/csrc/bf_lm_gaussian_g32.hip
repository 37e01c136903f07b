// bf_lm_gaussian_g32.hip — LM kernels for model id 0, 32 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 32
#include "bf_instantiate.hpp"
