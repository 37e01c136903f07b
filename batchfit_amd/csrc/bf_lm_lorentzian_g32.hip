// bf_lm_lorentzian_g32.hip — LM kernels for model id 1, 32 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 32
#include "bf_instantiate.hpp"
