// Command-line program `labels`: the reference's app/labels.cpp harness over
// the GPU path (variant KC_VARIANT_LABELS of kc_find_components).
#include "harness.h"

int main(int argc, char **argv) { return kc::RunApp(argc, argv, KC_VARIANT_LABELS, "labels"); }
