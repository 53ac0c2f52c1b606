/*
 * hddm_driver.c - the HDDM main loop of libfwhip, in C.
 *
 * Mirrors the control flow of the reference's parent solver in s-mode,
 * __HDDM_SOLVER_PARENT__ (lib/hddm/hddm_solver.h:59-183):
 *   - hddm_max_loop == 0: print "HDDM Solver was not executed!!!" and return
 *     without solving (hddm_solver.h:103-105);
 *   - LOOP_1ST + iteration_s_1st; status LAST at once when ||r0|| < 1e-60;
 *   - while (status != LAST) { LOOP + iteration_s; loop == max_loop ->
 *     status LAST (without replacing P by U, as the reference does) };
 *   - LOOP_LAST with the final interface values.
 * The steps themselves run on the device (fw_host.hip).
 */
#include <stdio.h>

#include "../../include/fwhip.h"

int fwh_solve(fwh_ctx *ctx, const fwh_opts *opts, fwh_stats *stats) {
  int rc, done = 0;
  if (!ctx || !opts) return FWH_E_ARG;
  if (opts->hddm_max_loop == 0) {
    fprintf(stdout, "HDDM Solver was not executed!!!\n");
    fflush(stdout);
    if (stats) {
      fwh_stats z = {0};
      *stats = z;
    }
    return FWH_OK;
  }
  rc = fwh_first(ctx, opts);
  if (rc) return rc;
  /* status after the first step */
  rc = fwh_iterate(ctx, opts, 0, &done);
  if (rc) return rc;
  while (!done) {
    int chunk = opts->graph_iters > 0 ? opts->graph_iters * 8 : 64;
    rc = fwh_iterate(ctx, opts, chunk, &done);
    if (rc) return rc;
  }
  rc = fwh_last(ctx);
  if (rc) return rc;
  if (stats) rc = fwh_get_stats(ctx, stats);
  return rc;
}
