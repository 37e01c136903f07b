// bf_lm_gaussian_g4.hip — LM kernels for model id 0, 4 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 4
#include "bf_instantiate.hpp"
