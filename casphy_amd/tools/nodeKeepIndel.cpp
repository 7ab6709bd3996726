// nodeKeepIndel <prefix> <cutoff>
// Drop-in for cu/nodeKeepIndel.cu (called by R/CasPhy_functions.R:2281,2666):
// u32 states shifted by +1 for node states and outgroups.
#include "node_keep_tool.hpp"

int main(int argc, char **argv) { return tool::node_keep_main(argc, argv, true); }
