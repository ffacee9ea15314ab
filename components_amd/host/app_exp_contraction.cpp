// Command-line program `exp_contraction`: the reference's app/exp_contraction.cpp harness over
// the GPU path (variant KC_VARIANT_EXP of kc_find_components).
#include "harness.h"

int main(int argc, char **argv) { return kc::RunApp(argc, argv, KC_VARIANT_EXP, "exp"); }
