/* Oracle: WENO5 horizontal/vertical fluxes (test infrastructure only).
 * hadv_weno5: src/adv_routines.F:204-430; vadv_weno5: :2359-2557. */
#include "cm1o_internal.h"
#include <stdio.h>
#include <stdlib.h>

void o_hadv_weno5(cm1o_model *M, int stag, const float *rru, const float *rrv,
                  float *dumx, float *dumy, const float *a, int pdef, double weps) {
  (void)M; (void)stag; (void)rru; (void)rrv; (void)dumx; (void)dumy; (void)a; (void)pdef; (void)weps;
  fprintf(stderr, "cm1o: WENO5 advection is not built yet\n");
  abort();
}

void o_vadv_weno5(cm1o_model *M, int stag, const float *rrw, float *dumz, const float *a,
                  int pdef, int pdefweno, double weps) {
  (void)M; (void)stag; (void)rrw; (void)dumz; (void)a; (void)pdef; (void)pdefweno; (void)weps;
  fprintf(stderr, "cm1o: WENO5 advection is not built yet\n");
  abort();
}
