/*
 * rk2_redi_hip.h -- the whole-model seam of the reference's RK2 benchmark
 * for the Redi kernel, backed by libredi_hip.so.
 *
 * The reference's RK2 driver (ideal_benchmark) links one self-contained
 * -Bsymbolic shared library per kernel and reaches it only through
 * argument-less bind(C) entries (ideal_benchmark/wrap/rk2_redi.F90:77-199,
 * ideal_benchmark/README.md:21-24,132-163).  librk2_redi_hip.so exports the
 * same entries with the same arguments, so the driver can load it in place
 * of the reference's librediDC.so; every stage runs the gfx950 kernels on
 * state that stays resident on the device between stages.
 */
#ifndef RK2_REDI_HIP_H
#define RK2_REDI_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* rk2_redi_init(nxp, nyp, nz) (rk2_redi.F90:77-137): builds the seam's
 * analytic state (nghost = 3, uniform 8 km x 11.1 km metrics, all-wet,
 * linear EOS, khtr = 100 m^2/s through the use_ext face diffusivities,
 * walls on all edges), uploads it and runs one redi_calc_coeffs. */
void rk2_redi_init(int nxp, int nyp, int nz);

/* rk2_redi_stage (rk2_redi.F90:139-143): redi_calc_coeffs then
 * redi_apply_flux_hoist with dt = 900 s.  Synchronous. */
void rk2_redi_stage(void);

/* rk2_redi_stage_unopt (rk2_redi.F90:145-149): redi_calc_coeffs then
 * redi_apply_flux.  Same bits as rk2_redi_stage.  Synchronous. */
void rk2_redi_stage_unopt(void);

/* rk2_redi_stage_cuda / rk2_redi_stage_cuda_unopt (rk2_redi.F90:151-196):
 * one flat launch (redi_opt_launch / redi_cuda_launch_) on the same device
 * arrays.  Same bits as rk2_redi_stage. */
void rk2_redi_stage_cuda(void);
void rk2_redi_stage_cuda_unopt(void);

/* rk2_redi_probe (rk2_redi.F90:198-202): min and max of tracer 1 (T*h) over
 * every cell, ghosts included. */
void rk2_redi_probe(double *vmin, double *vmax);

/* No counterpart in the reference: copies the seam's current h and the two
 * tracers (nx, ny, nz each, nx = nxp + 6) to host buffers, so a test can run
 * the same state through the oracle.  Returns 0, or -1 before init. */
int rk2_redi_export(double *h, double *t_htr, double *s_htr);

#ifdef __cplusplus
}
#endif
#endif /* RK2_REDI_HIP_H */
