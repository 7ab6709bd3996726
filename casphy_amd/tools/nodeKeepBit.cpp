// nodeKeepBit <prefix> <cutoff>
// Drop-in for cu/nodeKeepBit.cu (called by R/CasPhy_functions.R:2277,2662):
// bit rows of node states and outgroups.
#include "node_keep_tool.hpp"

int main(int argc, char **argv) { return tool::node_keep_main(argc, argv, false); }
