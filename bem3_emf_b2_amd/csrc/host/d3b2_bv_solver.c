/* d3b2_bv_solver object_settings_file output.dat
 * Command line of the reference's d3b2_src/d3b2_bv_solver.c: places the
 * objects, runs the multi-object boundary-value solve on the GPU and writes
 * the analysis data file (plus the .particles node list). */
#include "bem3_emf_b2.h"

int main(int argc, char **argv)
{
  MOBJ mo;
  mo_initialize(argc, argv, &mo);
  mo_print_data(&mo);
  mo_output_node_particles(argv[2], &mo);
  mo_solve_bv(&mo);
  mo_dat_write(argv[2], &mo);
  mo_finalize(&mo);
  return 0;
}
