// bf_lm_two_lorentzian_g4.hip — LM kernels for model id 3, 4 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 4
#include "bf_instantiate.hpp"
