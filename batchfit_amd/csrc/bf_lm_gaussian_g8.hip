// bf_lm_gaussian_g8.hip — LM kernels for model id 0, 8 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 8
#include "bf_instantiate.hpp"
