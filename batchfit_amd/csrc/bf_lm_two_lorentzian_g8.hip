// bf_lm_two_lorentzian_g8.hip — LM kernels for model id 3, 8 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 8
#include "bf_instantiate.hpp"
