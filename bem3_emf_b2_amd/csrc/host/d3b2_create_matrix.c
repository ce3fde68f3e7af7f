/* d3b2_create_matrix medium_file mesh_file output.obj [rv_x rv_y rv_z theta tr_x tr_y tr_z]
 * Command line of the reference's d3b2_src/d3b2_create_matrix.c: reads one
 * object, builds its boundary-integral system on the GPU and writes the
 * object file read by d3b2_bv_solver. */
#include "bem3_emf_b2.h"

int main(int argc, char **argv)
{
  DOMD md;
  read_domd2(argc, argv, &md);
  print_domd2(&md);
  initialize_domd2(&md);
  create_matrix(&md, argv[3]);
  finalize_domd2(&md);
  return 0;
}
