// bf_lm_two_lorentzian_g16.hip — LM kernels for model id 3, 16 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 16
#include "bf_instantiate.hpp"
