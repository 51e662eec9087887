/* Test infrastructure only (never shipped): adapts the reference tree's own
 * euler_5wave_2D Riemann solvers, development/rp_approaches/rpn2_euler_5wave.f
 * and rpt2_euler_5wave.f, built from their source by ref.mk, to the argument
 * lists that the reference's classic/step2.f90 and step2ds.f90 pass
 * (classic/flux2.f90:101,181,194).  Those two files keep an older list: rpn2
 * without maux, rpt2 without maux and with `ilr` where the newer one passes
 * imp.  gamma reaches them through the /cparam/ common block
 * (rpn2_euler_5wave.f:44, rpt2_euler_5wave.f:25); rpt2 reads the Roe averages
 * that rpn2 left in /comroe/ for the same slice, which is the order flux2
 * calls them in. */
extern struct { double gamma, gamma1; } cparam_;

extern void rpn2_(const int *ixy, const int *maxm, const int *meqn, const int *mwaves,
                  const int *mbc, const int *mx, const double *ql, const double *qr,
                  const double *auxl, const double *auxr, double *wave, double *s,
                  double *amdq, double *apdq);
extern void rpt2_(const int *ixy, const int *maxm, const int *meqn, const int *mwaves,
                  const int *mbc, const int *mx, const double *ql, const double *qr,
                  const double *aux1, const double *aux2, const double *aux3, const int *ilr,
                  const double *asdq, double *bmasdq, double *bpasdq);

void euler5_set_gamma(double gamma) {
  cparam_.gamma = gamma;
  cparam_.gamma1 = gamma - 1.0;
}

void euler5_rpn2_(const int *ixy, const int *maxm, const int *meqn, const int *mwaves,
                  const int *maux, const int *mbc, const int *mx, const double *ql,
                  const double *qr, const double *auxl, const double *auxr, double *wave,
                  double *s, double *amdq, double *apdq) {
  (void)maux;
  rpn2_(ixy, maxm, meqn, mwaves, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq);
}

void euler5_rpt2_(const int *ixy, const int *imp, const int *maxm, const int *meqn,
                  const int *mwaves, const int *maux, const int *mbc, const int *mx,
                  const double *ql, const double *qr, const double *aux1, const double *aux2,
                  const double *aux3, const double *asdq, double *bmasdq, double *bpasdq) {
  (void)maux;
  rpt2_(ixy, maxm, meqn, mwaves, mbc, mx, ql, qr, aux1, aux2, aux3, imp, asdq, bmasdq, bpasdq);
}
