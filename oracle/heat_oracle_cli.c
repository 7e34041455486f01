/* Command-line front end of the CPU oracle (test infrastructure only):
 * same argv and outputs as src/heat_solver_het.cu's main for one rank. */
#include "heat_oracle.h"
int main(int argc, char** argv) { return ho_main(argc, argv); }
