/* SeisCL_MPI: the reference's program name, on one MI355X GPU
 * (src/SeisCL_MPI.c:34-200). */
#include "seiscl_cli.h"

int main(int argc, char **argv) { return seis_cli_main(argc, argv); }
