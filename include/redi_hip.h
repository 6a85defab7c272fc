/*
 * redi_hip.h -- C ABI of the MI355X (gfx950) Redi neutral-diffusion step.
 *
 * Drop-in boundary for the reference's ocean_redi call surface.  Every entry
 * point takes plain pointers and sizes; array pointers are DEVICE pointers
 * (HBM) laid out exactly like the reference's Fortran arrays (column-major:
 * i fastest, then j, then k or the neutral-surface index s).  Streams are
 * passed as `void*` holding a hipStream_t (NULL = the null stream).  Unless
 * stated otherwise a call only enqueues work on its stream and returns.
 *
 * Return codes: REDI_OK (0) = work enqueued; REDI_SKIPPED (1) = one of the
 * reference's silent no-op guards fired and nothing was done; negative =
 * error, with a message from redi_last_error().
 */
#ifndef REDI_HIP_H
#define REDI_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define REDI_OK 0
#define REDI_SKIPPED 1
#define REDI_ERR_ARG (-1)
#define REDI_ERR_NZ (-2)
#define REDI_ERR_HIP (-3)
#define REDI_ERR_STATE (-4)
#define REDI_ERR_CONFIG (-5)

/* Largest supported layer count (the reference caps 2*nz+2 at its
 * compile-time NZ_STACK_MAX, redi/ocean_redi.F90:483-486). */
#define REDI_NZ_MAX 256

#define REDI_EOS_LINEAR 1
#define REDI_EOS_WRIGHT97 2
#define REDI_EOS_ROQUET_SPV 4

#define REDI_OBC_WALL 1

/* Precision forms of a redi_t (redi_init_form). */
#define REDI_FORM_FP64 0       /* the plain form: matches the reference bits */
#define REDI_FORM_FP32_STORE 1 /* fields, records and fluxes in FP32; density
                                  differences, neutral positions and the tracer-
                                  tendency accumulation in FP64 */

/* hgrid_t (redi/grid.F90:17-25) */
typedef struct {
    int nx_total, ny_total; /* incl. ghosts: n_phys + 2*nghost */
    int nx_phys, ny_phys;
    int nghost;
    int pad_;
    double dx, dy;
} redi_hgrid_t;

/* ocean_metrics_t (redi/ocean_metrics.F90:14-22); device pointers */
typedef struct {
    const double *dy_cu; /* (nx+1, ny) */
    const double *dx_cv; /* (nx, ny+1) */
    const double *idxCu; /* (nx+1, ny) */
    const double *idyCv; /* (nx, ny+1) */
    const double *areaT; /* (nx, ny)   */
    const double *wet_u; /* (nx+1, ny) */
    const double *wet_v; /* (nx, ny+1) */
} redi_metrics_t;

/* ocean_eos_t (redi/ocean_eos.F90:239-278) */
typedef struct {
    int variant; /* REDI_EOS_* */
    int ts_convention;
    double rho0, T_ref, S_ref, alpha_T, beta_S, p_ref;
} redi_eos_t;

/* multilayer_cgrid_state_t (redi/multilayer_cgrid_state.F90:18-28).
 * `tracers` is a HOST array of ntracers device pointers to hTr (nx,ny,nz). */
typedef struct {
    int nz_ml;
    int idx_temperature; /* 1-based; 0 = not registered */
    int idx_salinity;
    int ntracers;
    const double *h_layer; /* (nx, ny, nz), device */
    double *const *tracers;
} redi_state_t;

/* The same state with FP32 fields, for REDI_FORM_FP32_STORE. */
typedef struct {
    int nz_ml;
    int idx_temperature;
    int idx_salinity;
    int ntracers;
    const float *h_layer;
    float *const *tracers;
} redi_state32_t;

/* ocean_bc_state_t (redi/ocean_boundary_types.F90:14-22): per-edge tags,
 * REDI_OBC_WALL = no-normal-flow wall. */
typedef struct {
    int west, east, south, north;
} redi_bc_t;

/* Phase-A coefficient arrays, the public components uPoL..vhEff of
 * ocean_redi_t (redi/ocean_redi.F90:105-110), plus khtr_u/khtr_v
 * (:115-116).  Device pointers; Ko are int32; layouts (nx+1,ny,ns),
 * (nx+1,ny,ns-1), (nx,ny+1,ns), (nx,ny+1,ns-1), (nx+1,ny), (nx,ny+1). */
typedef struct {
    double *uPoL, *uPoR;
    int *uKoL, *uKoR;
    double *uhEff;
    double *vPoL, *vPoR;
    int *vKoL, *vKoR;
    double *vhEff;
    double *khtr_u, *khtr_v;
} redi_coeffs_t;

/* Opaque ocean_redi_t. */
typedef struct redi_s redi_t;

/* ocean_redi_init (redi/ocean_redi.F90:466-507).  nz_ml < 1 is taken as 1;
 * 2*nz+2 > 2*REDI_NZ_MAX+2 returns REDI_ERR_NZ (the reference's error stop).
 * Allocates the Phase-A state in HBM and fills it like the reference's
 * `allocate(..., source=0 / 1)`.  Synchronous. */
int redi_init(redi_t **self, const redi_hgrid_t *grid, int nz_ml);

/* redi_init for a precision form (REDI_FORM_*).  A state of one form takes
 * only that form's entries (redi_calc_coeffs / redi_apply_flux / redi_step
 * for REDI_FORM_FP64, the _fp32 entries for REDI_FORM_FP32_STORE).  This
 * and the _fp32 entries have no counterpart in the reference, whose state
 * is FP64 only. */
int redi_init_form(redi_t **self, const redi_hgrid_t *grid, int nz_ml, int form);

/* ocean_redi_destroy (redi/ocean_redi.F90:509-529).  Synchronous. */
int redi_destroy(redi_t *self);

/* The physics knobs of ocean_redi_t (:78-90): enable, khtr (m^2/s),
 * continuous (0 is rejected, as the reference rejects it at configure),
 * ref_pres (held, never read, as in the reference). */
int redi_configure(redi_t *self, int enable, double khtr, int continuous,
                   double ref_pres);

/* Write the Phase-A state into caller-owned device arrays in the
 * reference layout (uPoL..vhEff as after the last redi_calc_coeffs, the
 * reference's init values on faces Phase A never visits; khtr_u/khtr_v as
 * after the last redi_apply_flux; either khtr pointer may be NULL).  The
 * library keeps the state as one 32-byte record per face and neutral
 * surface, so this is a layout conversion enqueued on `stream`. */
int redi_export_coeffs(const redi_t *self, const redi_coeffs_t *dst,
                       void *stream);

/* Thread-tile shape (tx * ty lanes: 64, 128, 192 or 256) of the neutral
 * sweep: tiles of adjacent faces share columns in cache.  Default 64 x 1. */
int redi_set_tiles(redi_t *self, int tx, int ty);

/* Tracers handled per flux launch (1..4, default 4): more tracers share one
 * pass over the face records, fewer need fewer registers. */
int redi_set_tracer_group(redi_t *self, int n);

/* Storage order of the per-step tracer reconstruction: per column (0,
 * default) or layer-major (1).  Changes only the schedule, never the
 * result. */
int redi_set_recon_order(redi_t *self, int layer_major);

/* Optional FP64 output of the tracer tendency dTr * (1/areaT) (the term
 * redi_apply_flux adds to hTr, :1176-1180) of every later redi_apply_flux:
 * tend[t] is a device array (nx, ny, nz) for tracer t; ntracers = 0 turns
 * it off.  No counterpart in the reference (its dTr is a local array); it
 * lets the FP32-store form's tendency be compared before it is rounded
 * into the FP32 state. */
int redi_set_tendency_output(redi_t *self, int ntracers, double *const *tend);

/* Optional diagnostic output of the down-gradient guard's decisions
 * (redi/ocean_redi.F90:913-924) in every later redi_apply_flux: for tracer
 * t, face f and chunk c of 63 sublayers (sublayers 63c .. 63c+62), bit b of
 * u_bits[(t*nu + f)*nch + c] is set when sublayer 63c+b contributes
 * (hEff != 0) and the guard zeroed its flux; nu = (nx+1)*ny u faces (v_bits:
 * nv = nx*(ny+1)), nch = ceil((2*nz+1)/63), device arrays of uint64 sized
 * ntracers*nu*nch and ntracers*nv*nch.  Words of faces Phase B does not
 * compute (edges, walls, faces without a contributing sublayer) are not
 * written.  NULL, NULL turns it off.  No counterpart in the reference: it
 * lets the number of guard decisions that differ between two runs (the
 * precision forms, rounded inputs) be counted. */
int redi_set_guard_output(redi_t *self, uint64_t *u_bits, uint64_t *v_bits);

/* Opt-in neutral-slope taper (no counterpart in the reference, which does
 * not taper): in every later redi_calc_coeffs, a sublayer whose neutral
 * slope |S| = |zR - zL| / dx (z the sublayer's mid-depth in the right / left
 * column, dx from idxCu / idyCv) exceeds slope_max has its hEff scaled by
 * (slope_max / |S|)^2.  slope_max = 0 (the default) turns it off; every
 * parity result with the reference is for the default. */
int redi_set_slope_taper(redi_t *self, double slope_max);

/* Order in which the sweep's face tiles are numbered (and so handed to
 * the 8 XCDs in contiguous bands): along the rows (0, default) or down the
 * rows first (1), which puts vertical neighbours -- sharing columns -- on
 * one L2 one after the other.  Changes only the schedule, never the
 * result. */
int redi_set_tile_order(redi_t *self, int vertical);

/* Cells per run of the Phase-B flux kernels (1..4096; 0, the default, picks
 * 32, or 16 or 8 on domains of under 640k or 320k columns): one wavefront
 * walks a run of cells along a row (u faces) or up a column (v faces), face
 * by face, and adds each face's sublayer fluxes into its two cells in the
 * reference's order.  A run's two end faces are also computed by the
 * neighbouring runs.  Changes only the schedule, never the result. */
int redi_set_flux_segment(redi_t *self, int cells);

/* Layers per workgroup of the per-column tiled reconstruction (8, 16 or 32,
 * default 16; the workgroup takes 256 / kt columns).  Changes only the
 * schedule, never the result. */
int redi_set_recon_chunk(redi_t *self, int kt);

/* Faces of the last redi_calc_coeffs whose contributing sublayers are not
 * monotone in layer index (u and v faces).  Synchronous. */
int redi_nonmonotone_faces(const redi_t *self, long long *n_u, long long *n_v);

/* redi_calc_coeffs (redi/ocean_redi.F90:572-607): Phase A. */
int redi_calc_coeffs(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
                     const redi_eos_t *eos, redi_t *self,
                     const redi_state_t *ms, void *stream);

/* redi_apply_flux (redi/ocean_redi.F90:927-1002): Phase B for every tracer.
 * khtr_u_ext/khtr_v_ext: both non-NULL = the VarMix (use_ext) path, else the
 * scalar khtr.  bc NULL = all four edges are walls. */
int redi_apply_flux(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
                    redi_t *self, const redi_state_t *ms, double dt,
                    const double *khtr_u_ext, const double *khtr_v_ext,
                    const redi_bc_t *bc, void *stream);

/* redi_calc_coeffs / redi_apply_flux / redi_step of the FP32-store form:
 * same contracts, FP32 h and tracers (the metrics, khtr_*_ext and dt stay
 * FP64).  Results are within the form's stated error of the FP64 form. */
int redi_calc_coeffs_fp32(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
                          const redi_eos_t *eos, redi_t *self,
                          const redi_state32_t *ms, void *stream);
int redi_apply_flux_fp32(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
                         redi_t *self, const redi_state32_t *ms, double dt,
                         const double *khtr_u_ext, const double *khtr_v_ext,
                         const redi_bc_t *bc, void *stream);
int redi_step_fp32(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
                   const redi_eos_t *eos, redi_t *self, const redi_state32_t *ms,
                   double dt, const double *khtr_u_ext, const double *khtr_v_ext,
                   const redi_bc_t *bc, int nsteps, void *stream);

/* redi_apply_flux_hoist (redi/ocean_redi.F90:1201-1262): same contract and
 * same bits as redi_apply_flux; here both run the same kernels, since the
 * column reconstruction is always done once per column. */
int redi_apply_flux_hoist(const redi_hgrid_t *grid,
                          const redi_metrics_t *metrics, redi_t *self,
                          const redi_state_t *ms, double dt,
                          const double *khtr_u_ext, const double *khtr_v_ext,
                          const redi_bc_t *bc, void *stream);

/* One full step: redi_calc_coeffs then redi_apply_flux, repeated nsteps
 * times (Phase B is not idempotent; hTr accumulates). */
int redi_step(const redi_hgrid_t *grid, const redi_metrics_t *metrics,
              const redi_eos_t *eos, redi_t *self, const redi_state_t *ms,
              double dt, const double *khtr_u_ext, const double *khtr_v_ext,
              const redi_bc_t *bc, int nsteps, void *stream);

/* Bytes of device scratch an ocean_redi_t of this size holds for the
 * column state / reconstruction buffers, and the total device bytes it
 * holds (scratch plus the Phase-A face state). */
long long redi_workspace_bytes(int nx_total, int ny_total, int nz);
long long redi_state_bytes(int nx_total, int ny_total, int nz);
long long redi_state_bytes_form(int nx_total, int ny_total, int nz, int form);

/* ---- per-kernel timing ---------------------------------------------------
 * When enabled, every kernel launch made for this ocean_redi_t is bracketed
 * by HIP events recorded on the launch's own stream.  redi_timing_collect
 * waits for those events, adds their elapsed times per kernel kind into
 * `out` and forgets them. */
#define REDI_NKERNELS 9
#define REDI_K_COLUMN_STATE 0
#define REDI_K_SWEEP_U 1
#define REDI_K_SWEEP_V 2
#define REDI_K_FACE_KHTR 3
#define REDI_K_TRACER_RECON 4
#define REDI_K_FLUX_X 5         /* u-face fluxes, added as W and E contributions */
#define REDI_K_FLUX_Y 6         /* v-face fluxes, added as S and N contributions */
#define REDI_K_TRACER_UPDATE 7  /* hTr += dTr / areaT */
#define REDI_K_SWEEP_UV 8       /* both face grids' sweeps in one launch (64-lane tiles) */
typedef struct {
    double ms[REDI_NKERNELS];           /* summed kernel time */
    long long launches[REDI_NKERNELS];  /* launches timed */
} redi_kernel_times_t;
int redi_timing_enable(redi_t *self, int on);
int redi_timing_collect(redi_t *self, redi_kernel_times_t *out);
const char *redi_kernel_name(int kind);

/* Message of the last error on this host thread ("" if none). */
const char *redi_last_error(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int redi_version(void);

/* ---- flat compatibility entries ------------------------------------------
 * Same symbol, argument list and semantics as the reference's hand-written
 * GPU launchers redi_opt_launch (redi/opt_kernel.cu:534-574) and
 * redi_cuda_launch_ (redi/redi_kernel.cu:537-580): Fortran by-reference
 * scalars, caller-owned device arrays, exactly two tracers (t_htr, s_htr,
 * updated in place), linear EOS, use_ext path, walls ww/we/wsz/wn (1 =
 * wall), default stream, synchronous (ends with a device synchronise),
 * no return code.  tr_snap is accepted and not used (the reconstruction
 * buffers are the snapshot).  Scratch beyond the caller's arrays comes from
 * a per-device grow-only cache. */
void redi_opt_launch(
    const int *nx, const int *ny, const int *nz, const int *ns,
    const double *dt, const int *nghost, const int *nxp, const int *nyp,
    const int *ww, const int *we, const int *wsz, const int *wn,
    const double *rho0, const double *T_ref, const double *S_ref,
    const double *alpha_T, const double *beta_S, const double *h_layer,
    double *t_htr, double *s_htr, const double *wet_u, const double *wet_v,
    double *uPoL, double *uPoR, int *uKoL, int *uKoR, double *uhEff,
    double *vPoL, double *vPoR, int *vKoL, int *vKoR, double *vhEff,
    const double *khtr_u_ext, const double *khtr_v_ext, double *khtr_u,
    double *khtr_v, const double *dy_cu, const double *dx_cv,
    const double *idxCu, const double *idyCv, const double *areaT,
    double *tr_snap);
void redi_cuda_launch_(
    const int *nx, const int *ny, const int *nz, const int *ns,
    const double *dt, const int *nghost, const int *nxp, const int *nyp,
    const int *ww, const int *we, const int *wsz, const int *wn,
    const double *rho0, const double *T_ref, const double *S_ref,
    const double *alpha_T, const double *beta_S, const double *h_layer,
    double *t_htr, double *s_htr, const double *wet_u, const double *wet_v,
    double *uPoL, double *uPoR, int *uKoL, int *uKoR, double *uhEff,
    double *vPoL, double *vPoR, int *vKoL, int *vKoR, double *vhEff,
    const double *khtr_u_ext, const double *khtr_v_ext, double *khtr_u,
    double *khtr_v, const double *dy_cu, const double *dx_cv,
    const double *idxCu, const double *idyCv, const double *areaT,
    double *tr_snap);

/* Extension of the flat entry to any tracer count and EOS variant, with a
 * return code and a stream (asynchronous): the nt tracer device pointers are
 * passed in the host array `tracers`; idx_T/idx_S are 1-based. */
int redi_launch_nt(int nx, int ny, int nz, double dt, int nghost, int nxp,
                   int nyp, int ww, int we, int wsz, int wn,
                   const redi_eos_t *eos, const double *h_layer, int nt,
                   double *const *tracers, int idx_T, int idx_S,
                   const double *wet_u, const double *wet_v,
                   const redi_coeffs_t *coeffs, const double *khtr_u_ext,
                   const double *khtr_v_ext, const double *dy_cu,
                   const double *dx_cv, const double *idxCu,
                   const double *idyCv, const double *areaT, void *stream);

/* redi_launch_nt without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * redi_launch_nt keeps its face state and workspace in a per-device cache
 * that it hipMallocs on first use and grows; this entry keeps them in one
 * static device array of REDI_NOALLOC_SCRATCH_BYTES (1 GiB) instead,
 * allocated with the library's code on each device, and calls no hipMalloc,
 * hipFree or allocating routine.  Domains whose face state and workspace
 * need more (about 216 B per cell at 2 tracers: 473x297x30 plus ghosts
 * fits) return REDI_ERR_BUDGET.  Every call on one device shares its array:
 * calls that overlap on a device race.  Same arguments, results and
 * stream semantics as redi_launch_nt. */
int redi_launch_nt_noalloc(int nx, int ny, int nz, double dt, int nghost,
                           int nxp, int nyp, int ww, int we, int wsz, int wn,
                           const redi_eos_t *eos, const double *h_layer,
                           int nt, double *const *tracers, int idx_T,
                           int idx_S, const double *wet_u,
                           const double *wet_v, const redi_coeffs_t *coeffs,
                           const double *khtr_u_ext, const double *khtr_v_ext,
                           const double *dy_cu, const double *dx_cv,
                           const double *idxCu, const double *idyCv,
                           const double *areaT, void *stream);
#define REDI_NOALLOC_SCRATCH_BYTES (1ull << 30)

/* ---- batched form: many independent domains in one launch per kernel -----
 * No counterpart in the reference, which steps one domain per call.  A batch
 * holds up to max_domains independent domains of any size, each with its own
 * grid, layer count, EOS, walls, dt and diffusivity, all in one precision
 * form.  redi_batch_step advances every domain in the batch by one Redi step
 * (redi_calc_coeffs then redi_apply_flux, as redi_step) with one hipGraph
 * launch: one persistent kernel per kernel kind walks a device-resident work
 * table that covers every domain.  Each domain's result is bit-identical to
 * redi_step on that domain alone, in the same form.
 *
 * The Phase-A state and scratch of every domain are carved from one device
 * arena of arena_bytes allocated at redi_batch_create; adding a domain
 * allocates nothing on the device, and a domain that does not fit returns
 * REDI_ERR_BUDGET.  The caller owns h, the tracers and the metrics of each
 * domain (device pointers, which must stay valid while the domain is in the
 * batch).  All calls on one batch must be made from one host thread and on
 * one stream (or the caller orders the streams). */
#define REDI_ERR_BUDGET (-6)

typedef struct redi_batch_s redi_batch_t;

typedef struct {
    redi_hgrid_t grid;
    int nz_ml;
    int ntracers;         /* 1..group of the batch */
    int idx_temperature;  /* 1-based */
    int idx_salinity;
    redi_metrics_t metrics;
    redi_eos_t eos;
    const void *h_layer;  /* (nx, ny, nz), device, the batch form's type */
    void *const *tracers; /* host array of ntracers device pointers (nx, ny, nz) */
    double dt;
    double khtr;          /* used where khtr_u_ext / khtr_v_ext are NULL; must be > 0 then */
    const double *khtr_u_ext, *khtr_v_ext; /* both non-NULL = the use_ext path */
    redi_bc_t bc;
    int tile_i, tile_j;   /* sweep tile, tile_i * tile_j == 64 (0, 0 = 64 x 1) */
    int kchunk;           /* layers per reconstruction workgroup: 8, 16, 32 (0 = 16) */
    int pad_;
    double slope_max;     /* opt-in slope taper (redi_set_slope_taper), 0 = off */
} redi_domain_t;

/* group: tracer slots of the batch's flux kernels (2..4); every domain has at
 * most that many tracers.  persist: workgroups of each persistent kernel per
 * CU per occupancy slot (1 = exactly fill the GPU once; 0 = one workgroup per
 * work block, no persistence). */
int redi_batch_create(redi_batch_t **out, int form, int group, int max_domains,
                      long long arena_bytes, int persist);
int redi_batch_destroy(redi_batch_t *b);

/* Arena bytes one domain of this size takes in a batch of this form / group. */
long long redi_batch_domain_bytes(int nx_total, int ny_total, int nz, int form, int group);

/* Admit a domain: returns its slot (>= 0) or a negative error; enqueues the
 * initial fill of its Phase-A state on `stream`. */
int redi_batch_add(redi_batch_t *b, const redi_domain_t *d, void *stream);
/* Retire a slot; its arena range is reused by later adds (in stream order). */
int redi_batch_remove(redi_batch_t *b, int slot);

/* nsteps Redi steps of every domain in the batch, one graph launch per step. */
int redi_batch_step(redi_batch_t *b, int nsteps, void *stream);
/* The same steps as one launch sequence per domain (the unbatched schedule,
 * for comparison); same bits. */
int redi_batch_step_serial(redi_batch_t *b, int nsteps, void *stream);

/* Phase-A state of one slot in the reference layout (as redi_export_coeffs). */
int redi_batch_export_coeffs(const redi_batch_t *b, int slot, const redi_coeffs_t *dst,
                             void *stream);

/* Domains in the batch, arena bytes in use and free (largest free range). */
int redi_batch_info(const redi_batch_t *b, int *ndomains, long long *used_bytes,
                    long long *largest_free_bytes);

#ifdef __cplusplus
}
#endif
#endif /* REDI_HIP_H */
