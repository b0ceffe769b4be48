// Executable wrapper for the engine-app compatible command line.
#include "epirust_gpu.h"

int main(int argc, char** argv) { return epi_cli_main(argc, argv); }
