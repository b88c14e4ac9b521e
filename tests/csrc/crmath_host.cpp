// Host build of the renderer's correctly rounded acos / cos / pow (diffvg_amd/csrc/
// dvg_crmath.h) for the CPU tests, which compare it with the host libm and with
// high-precision values.  Test infrastructure only.
#include "../../diffvg_amd/csrc/dvg_crmath.h"

extern "C" {
void crm_eval(int fn, const double *x, double *y, int n) {
    for (int i = 0; i < n; i++) {
        if (fn == 0)
            y[i] = dvg::cr::cos_cr(x[i]);
        else if (fn == 1)
            y[i] = dvg::cr::acos_cr(x[i]);
        else
            y[i] = dvg::cr::pow_cr(x[i], 1. / 3.);
    }
}
}
