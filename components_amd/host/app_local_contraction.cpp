// Command-line program `local_contraction`: the reference's app/local_contraction.cpp harness over
// the GPU path (variant KC_VARIANT_LOCAL of kc_find_components).
#include "harness.h"

int main(int argc, char **argv) { return kc::RunApp(argc, argv, KC_VARIANT_LOCAL, "local"); }
