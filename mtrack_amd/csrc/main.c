/*
 * mbtrack-mi355x: command-line drop-in for mbtrack-mpi (src/mbtrack-mpi.c:23-162).
 *
 *   mbtrack-mi355x input.conf [work_dir]        (work_dir defaults to "work/")
 *
 * One process on one GPU (device from MBT_DEVICE, default 0); there is no
 * rank-count rule.  BunchModel WEAK runs the weak-weak manager, STRONG the
 * weak-strong one (src/mbtrack-mpi.c:131-156).  Returns 0, or -1 on error like
 * the reference.
 */
#include <stdio.h>
#include <stdlib.h>
#include "mbtrack_mi355x.h"

int main(int argc, char **argv)
{
  const char *work = "work/";
  if (argc == 3) work = argv[2];
  else if (argc != 2) {
    fprintf(stderr, "ERROR ! \n\nUsage:\nmbtrack-mi355x inputfile.conf [/path/to/work]\n");
    return -1;
  }
  static mbt_config cfg;
  if (mbt_config_read(argv[1], work, &cfg) != MBT_OK) {
    fprintf(stderr, "ERROR: read_input failed (%s)\n", mbt_last_error());
    return -1;
  }
  const char *dev = getenv("MBT_DEVICE");
  int rc;
  if (cfg.bunch_model == 0) rc = mbt_manager_weakweak(&cfg, dev ? atoi(dev) : 0);
  else {
    /* BunchModel = STRONG: weak-strong + FBII; MBT_WS_BUNCH_CURRENTS=1 gives the
       bunches the filling's currents (the reference leaves them zero) */
    const char *cur = getenv("MBT_WS_BUNCH_CURRENTS");
    rc = mbt_manager_weakstrong(&cfg, dev ? atoi(dev) : 0, cur && atoi(cur) ? MBT_WS_BUNCH_CURRENTS : 0);
  }
  if (rc != MBT_OK) fprintf(stderr, "ERROR: %s\n", mbt_last_error());
  mbt_config_free(&cfg);
  return rc == MBT_OK ? 0 : -1;
}
