// main.cpp -- the `vmachine` executable (ref src/main.cpp:3-16); the command
// line lives in libvmc_host.so so that it can be driven in-process too.
#include "vmc_host.h"

int main(int argc, char** argv) { return vmc_host_main(argc, argv); }
