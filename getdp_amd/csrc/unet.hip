// Learned preconditioner (placeholder until the U-Net kernels land).
#include "krylov.h"

namespace gdp {

struct UserPrec {
  int n = 0;
};

int user_prec_apply(UserPrec *u, const double *in, double *out, KCtl *C, int gmode)
{
  (void)u;
  (void)in;
  (void)out;
  (void)C;
  (void)gmode;
  return set_error(GDP_E_UNSUPPORTED, "no learned preconditioner attached");
}

}  // namespace gdp
