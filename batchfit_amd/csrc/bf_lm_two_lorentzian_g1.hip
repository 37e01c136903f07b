// bf_lm_two_lorentzian_g1.hip — LM kernels for model id 3, 1 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 1
#include "bf_instantiate.hpp"
