// Command-line program `shortcuts`: the reference's app/shortcuts.cpp harness over
// the GPU path (variant KC_VARIANT_SHORTCUT of kc_find_components).
#include "harness.h"

int main(int argc, char **argv) { return kc::RunApp(argc, argv, KC_VARIANT_SHORTCUT, "shortcuts"); }
