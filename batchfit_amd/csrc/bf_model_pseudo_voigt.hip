// bf_model_pseudo_voigt.hip — kernel instantiations and launchers for model id 2.
#define BF_MODEL 2
#include "bf_instantiate.hpp"
