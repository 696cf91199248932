// Host build of the brick rules the device runs (brick_rules.hpp), for the tests.
#include "brick_rules.hpp"

extern "C" long icam_rules_bridge(double* val, uint8_t* st, double h, double clamp) { return icam::bridge_nodes(val, st, h, clamp); }
