// bf_model_two_lorentzian.hip — kernel instantiations and launchers for model id 3.
#define BF_MODEL 3
#include "bf_instantiate.hpp"
