// dispatch.hip -- routes each launcher of the kernel sources to its arithmetic
// variant (variant.hpp): CLAWHIP_F64 to claw::exact (no contraction: the
// reference's FP64 rounding), CLAWHIP_F32K and CLAWHIP_F64_FAST to claw::fast
// (FMA-contracted).  Launchers without floating-point work of their own (flag
// lists, ensemble tables, the FP64-only dimension sweeps of fmod) take exact.
#include "launch.hpp"

namespace claw {

namespace {
inline bool fast_dtype(int dtype) { return dtype != CLAWHIP_F64; }
}  // namespace

int launch_unsplit(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                   void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                   const int* lo, const int* hi, const int* refl, int ng, const int* method,
                   const int* mthlim, double dx, double dy, double dt,
                   unsigned long long* cfl_bits, hipStream_t stream, const AuxIn& aux) {
  return (fast_dtype(dtype) ? fast::launch_unsplit : exact::launch_unsplit)(
      rp, dtype, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl, ng, method, mthlim, dx,
      dy, dt, cfl_bits, stream, aux);
}

int launch_split2(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dt, const AuxIn& aux, unsigned long long* cfl, hipStream_t s) {
  return (fast_dtype(dtype) ? fast::launch_split2 : exact::launch_split2)(
      rp, dtype, q_in, gin, q_out, gout, comp_in, comp_out, work, cwork, lo, hi, refl, ng, method,
      mthlim, dx, dy, dt, aux, cfl, s);
}

int launch_split3(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dz, double dt, unsigned long long* cfl, hipStream_t s) {
  return (fast_dtype(dtype) ? fast::launch_split3 : exact::launch_split3)(
      rp, dtype, q_in, gin, q_out, gout, comp_in, comp_out, work, cwork, lo, hi, refl, ng, method,
      mthlim, dx, dy, dz, dt, cfl, s);
}

int launch_sharp_ssp33_run(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                           const int* lo, const int* hi, const int* refl, double dx, double dy,
                           const SharpRunBufs& b, SharpCtl* ctl, int nsteps,
                           hipStream_t stream) {
  return (fast_dtype(dtype) ? fast::launch_sharp_ssp33_run : exact::launch_sharp_ssp33_run)(
      rp, dtype, char_decomp, g, lo, hi, refl, dx, dy, b, ctl, nsteps, stream);
}

int launch_flag_list(const unsigned char* flags, int mx, int my, int* list, int* count,
                     hipStream_t s, const int* gate, bool zeroed) {
  return exact::launch_flag_list(flags, mx, my, list, count, s, gate, zeroed);
}

int launch_sharp_stage(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                       const Geom& gin, const int* lo, const int* hi, const int* refl, int ymode,
                       int omode, double dt, double s, double dx, double dy, const void* P,
                       const void* Q, const void* A, const void* B, const void* aux_in,
                       const void* comp_in, void* out, void* aux_out, void* comp_out,
                       unsigned long long* cfl, const FlagSet& flags, hipStream_t stream,
                       const AuxIn& aux, const void* src) {
  return (fast_dtype(dtype) ? fast::launch_sharp_stage : exact::launch_sharp_stage)(
      rp, dtype, char_decomp, g, gin, lo, hi, refl, ymode, omode, dt, s, dx, dy, P, Q, A, B,
      aux_in, comp_in, out, aux_out, comp_out, cfl, flags, stream, aux, src);
}

int launch_ssp104_mid(int dtype, void* q, void* s1, long long n, hipStream_t stream) {
  return (fast_dtype(dtype) ? fast::launch_ssp104_mid : exact::launch_ssp104_mid)(dtype, q, s1,
                                                                                 n, stream);
}

int launch_lmm_update(int dtype, int mode, long long n, const void* q, const void* u,
                      const void* R, const void* v, const double* c, void* out,
                      hipStream_t stream) {
  return (fast_dtype(dtype) ? fast::launch_lmm_update : exact::launch_lmm_update)(
      dtype, mode, n, q, u, R, v, c, out, stream);
}

int launch_unsplit_ctl(const clawhip_rp& rp, int dtype, const Geom& g, const int* lo,
                       const int* hi, const int* refl, int ng, const int* method,
                       const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                       hipStream_t stream) {
  return (fast_dtype(dtype) ? fast::launch_unsplit_ctl : exact::launch_unsplit_ctl)(
      rp, dtype, g, lo, hi, refl, ng, method, mthlim, ctl, bufs, nsteps, stream);
}

long long ensemble_table_bytes(int n) { return exact::ensemble_table_bytes(n); }

int ensemble_prepare(const clawhip_rp& rp, const clawhip_ens_member* members, int n,
                     const int* lo, const int* hi, const int* refl, int ng, const int* method,
                     const int* mthlim, int rows, void* table_dev, long long* waves) {
  return exact::ensemble_prepare(rp, members, n, lo, hi, refl, ng, method, mthlim, rows,
                                 table_dev, waves);
}

int ensemble_run(const clawhip_rp& rp, int dtype, const void* table_dev, const long long* waves,
                 int nsteps, hipStream_t s) {
  return (fast_dtype(dtype) ? fast::ensemble_run : exact::ensemble_run)(rp, dtype, table_dev,
                                                                       waves, nsteps, s);
}

int launch_sweep2(const clawhip_rp& rp, int ids, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  const AuxIn& aux, unsigned long long* cfl, hipStream_t s) {
  return exact::launch_sweep2(rp, ids, q_in, gin, q_out, gout, ng, method, mthlim, dtdx, aux, cfl,
                              s);
}

int launch_sweep3(const clawhip_rp& rp, int idir, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  unsigned long long* cfl, hipStream_t s) {
  return exact::launch_sweep3(rp, idir, q_in, gin, q_out, gout, ng, method, mthlim, dtdx, cfl, s);
}

}  // namespace claw
