// Command-line program `all_reduce`: the reference's app/all_reduce.cpp harness
// over the GPU path.  AllReduce (lib/components/all_reduce.h:46-158) gathers
// the whole graph on the root PE and labels it there sequentially; on one GPU
// the whole graph is already in one place, so every edge goes to the base case
// (min-label propagation with pointer jumping) and no contraction level runs.
// Labels are the canonical minimum ids (the reference labels each component
// by its BFS start in gather order).
#include "harness.h"

int main(int argc, char **argv) { return kc::RunApp(argc, argv, KC_VARIANT_EXP, "allreduce", true); }
