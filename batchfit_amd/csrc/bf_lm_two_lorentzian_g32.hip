// bf_lm_two_lorentzian_g32.hip — LM kernels for model id 3, 32 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 32
#include "bf_instantiate.hpp"
