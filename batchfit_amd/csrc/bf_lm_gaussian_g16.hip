// bf_lm_gaussian_g16.hip — LM kernels for model id 0, 16 lane(s) per fit.
#define BF_MODEL 0
#define BF_GROUP 16
#include "bf_instantiate.hpp"
