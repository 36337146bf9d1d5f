/*
 * plasmasim.h - C-ABI of the MI355X-native 3D electromagnetic PIC time step.
 *
 * The reference program (ashida-lab/PLASMAsim, PIC/3d_pic458_domain.c) has no
 * plugin layer: its hot path is the set of C routines that fdtd() calls on
 * global arrays (:627-952).  Each entry point below replaces one of those
 * routines, keeps its position in the step, and operates on device-resident
 * state owned by a pic_ctx instead of the reference's globals.
 *
 * Conventions
 *   - Axes and indices follow the reference: k (x), l (y), m (z), m contiguous,
 *     2 guard cells per side, so field arrays are (nx+4)*(ny+4)*(nz+4) and
 *     element (k,l,m) lives at ((k*(ny+4))+l)*(nz+4)+m.  z is the flow axis
 *     (reference: solar wind along +z, :425); x and y are periodic.
 *   - Particle coordinates are in length units with the +2 guard offset, as
 *     in the reference (interior cell k covers [(k+2)dx,(k+3)dx)).
 *   - Staggering (reference shape_func_v2 :3104-3139, external_current7
 *     :4579-4634): Ex,Jx at (k, l+1/2, m+1/2); Ey,Jy at (k+1/2, l, m+1/2);
 *     Ez,Jz at (k+1/2, l+1/2, m); Bx at (k+1/2, l, m); By at (k, l+1/2, m);
 *     Bz at (k, l, m+1/2); ni, ne at (k+1/2, l+1/2, m+1/2).
 *   - Every call returns 0 or a negative PIC_E_* code and never exits the
 *     process; pic_last_error() gives the message.  Calls are asynchronous on
 *     the context's HIP stream; downloads and diagnostics synchronize.
 *   - One context per host thread.  The library holds no global state.
 */
#ifndef PLASMASIM_H
#define PLASMASIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_ABI_VERSION 1

/* status codes */
#define PIC_OK 0
#define PIC_E_ARG (-1)      /* invalid argument / config */
#define PIC_E_NOMEM (-2)    /* device or host allocation failed */
#define PIC_E_HIP (-3)      /* HIP runtime error */
#define PIC_E_CAPACITY (-4) /* particle store would overflow its capacity */
#define PIC_E_STATE (-5)    /* call not valid in the current state */
#define PIC_E_IO (-6)       /* file could not be opened, read or written */
#define PIC_E_FORMAT (-7)   /* checkpoint header, config or checksum mismatch */

/* numeric forms */
#define PIC_FP64 0 /* plain form: reference FP64 arithmetic */
#define PIC_FP32 1 /* chosen form: int32 cell + fp32 offset, fp32 fields */

/* deposit modes */
#define PIC_DEPOSIT_ATOMIC 0        /* LDS pre-reduction + floating atomics */
#define PIC_DEPOSIT_DETERMINISTIC 1 /* fixed-point LDS + integer atomics:
                                       bit-reproducible run to run */

/* species */
#define PIC_ION 0
#define PIC_ELECTRON 1

/* field planes (upload/download selectors) */
enum pic_field {
  PIC_EX = 0, PIC_EY, PIC_EZ,
  PIC_BX, PIC_BY, PIC_BZ,          /* total B, as the reference stores it */
  PIC_B0X, PIC_B0Y, PIC_B0Z,       /* static dipole + IMF */
  PIC_JIX, PIC_JIY, PIC_JIZ,       /* ion current (deposited with +q) */
  PIC_JEX, PIC_JEY, PIC_JEZ,       /* electron current (deposited with +q) */
  PIC_NI, PIC_NE,                  /* number densities */
  PIC_NFIELDS
};
/* download only: the potential of the last Poisson correction in the
 * current step (zero when poisson_every is 0 or renew_grid ran since) */
#define PIC_PHI PIC_NFIELDS

/*
 * Runtime replacement of the reference's compile-time configuration
 * (:50-108), in normalized units (c = eps0 = mu0 = 1 for the shipped
 * presets).  Fill with pic_config_preset() and adjust.
 */
typedef struct pic_config {
  int32_t nx, ny, nz;   /* interior cells (reference Grid_Nx/Ny/Nz) */
  int32_t ppc;          /* macro-particles per cell per species (Np) */
  double dx, dt;
  double c, eps0, mu0;
  double q;             /* elementary charge; both species deposit with +q */
  double mass[2];       /* ion, electron (positive; the push uses -m_e) */
  double weight[2];     /* density carried by one macro-particle (p->n) */
  double vth[2];        /* per-component Maxwellian sigma, sqrt(kT/m) */
  double vdrift[3];     /* mean inflow velocity (reference (0,0,V)) */
  double imf[3];        /* uniform interplanetary field added to B0 */
  double dipole_m;      /* dipole moment magnitude (reference I*pi*R^2) */
  double dipole_alpha;  /* moment direction (0, sin a, cos a) (:4601-4619) */
  double dipole_soft;   /* inside this radius (cells) B0 is clamped to its
                           value at the softening radius; 0 = none */
  double dipole_rmax;   /* B0 = IMF beyond this radius in cells (210) */
  int32_t absorb[3];    /* sigma-layer width in cells per axis (0 = none) */
  double sigma_amp;     /* sigma profile amplitude (reference a, :304) */
  double ballistic_lo[3]; /* particles below/above these coordinates are */
  double ballistic_hi[3]; /* drifted without field push (:642-647, :735)  */
  int32_t periodic[2];  /* x, y: 1 = periodic wrap, 0 = particles dropped */
  int32_t inflow_every; /* add_*_flat cadence in steps (reference 5) */
  int32_t sort_every;   /* sorting cadence in steps (reference 500) */
  int32_t refill_far;   /* 1 = refill the downstream z slab too (reference) */
  int32_t precision;    /* PIC_FP64 or PIC_FP32 */
  int32_t deposit;      /* PIC_DEPOSIT_ATOMIC or PIC_DEPOSIT_DETERMINISTIC */
  int32_t tile[3];      /* sort tiles in cells; the sort key is tile-major,
                           lexicographic (k,l,m) inside a tile.  A tile equal
                           to the whole interior (or 0) gives exactly the
                           reference's key ((k*Ny)+l)*Nz+m (:7488-7497). */
  uint64_t seed;        /* Philox4x32-10 key for loading and inflow */
  int64_t capacity;     /* per-species particle capacity; 0 = automatic */
  int32_t poisson_every; /* renew_phi3 + sr_phi + E correction (:901-906)
                            every this many steps, after the E step; 0 =
                            off, as in the reference (PHI() is empty, :107).
                            Needs power-of-two nx, ny, nz. */
  int32_t phi_margin;   /* cells of the Poisson source's margin (reference
                           Absorb_grid3, :97): the source is zero unless
                           phi_margin < k-2 < nx-phi_margin (same for l) and
                           phi_margin <= m-2 < nz-phi_margin (:8340-8344) */
} pic_config;

/* Diagnostics (reference thrust_f :4940-5122, extended; all FP64). */
typedef struct pic_diag {
  double field_energy_e;     /* 1/2 eps0 sum |E|^2 dx^3, interior */
  double field_energy_b;     /* 1/2 sum |B|^2 / mu0 dx^3 (total B) */
  double field_energy_b1;    /* same for the induced field B - B0 */
  double kinetic[2];         /* sum (gamma-1) m c^2 w dx^3 per species */
  double kinetic_nr[2];      /* reference form 1/2 m w dx^3 v^2 */
  double momentum[2][3];     /* sum gamma m v w dx^3 per species */
  double charge;             /* sum q (ni - ne) dx^3 over the interior */
  double gauss_l2;           /* sqrt(sum r^2) of r = div E - q(ni-ne)/eps0 */
  double gauss_max;          /* max |r| over the interior */
  double force[15];          /* reference F[0..14] over the central
                                20-80 % box (ions flag>0, electrons) */
  double energy_box[6];      /* reference E[0..5] over the same box */
  int64_t count[2];          /* live particles per species */
} pic_diag;

typedef struct pic_ctx pic_ctx;

/* ---- configuration and lifetime ---- */

/* sizeof(pic_config) and sizeof(pic_diag) as compiled into the library, for
 * bindings to check their struct layout. */
size_t pic_sizeof_config(void);
size_t pic_sizeof_diag(void);

/* Fill *cfg with the workload presets: 1 = 32^3 x 16 ppc (FP64),
 * 2 = 128^3 x 64 ppc, 4 = 64^3 x 32 ppc, 5 = 256^3 x 32 ppc, m_i/m_e = 400;
 * 0 = a tiny 8x8x16 grid for tests.  Replaces the #define block :50-108. */
int pic_config_preset(int preset, pic_config *cfg);

/* Allocate device state and set up static fields.  Replaces init_grid
 * (:291-410) plus external_current7 (:4579-4634): E = 0, B = B0 = dipole,
 * sigma profile.  No particles are loaded. */
int pic_create(const pic_config *cfg, pic_ctx **out);
void pic_destroy(pic_ctx *ctx);
const char *pic_last_error(const pic_ctx *ctx);
int pic_get_config(const pic_ctx *ctx, pic_config *cfg);
void *pic_stream(pic_ctx *ctx); /* the context's hipStream_t */
int pic_sync(pic_ctx *ctx);

/* ---- data movement (host <-> device) ---- */

/* Host buffers hold (nx+4)*(ny+4)*(nz+4) doubles in the reference order. */
int pic_upload_field(pic_ctx *ctx, int which, const double *host);
int pic_download_field(pic_ctx *ctx, int which, double *host);

/* Replace the species' particles with n particles (FP64 SoA). */
int pic_upload_particles(pic_ctx *ctx, int species, int64_t n,
                         const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz);
/* Copy the live particles, in storage order, into caller arrays of length
 * >= cap; *n receives the live count (PIC_E_CAPACITY if cap is short). */
int pic_download_particles(pic_ctx *ctx, int species, int64_t cap, int64_t *n,
                           double *x, double *y, double *z,
                           double *vx, double *vy, double *vz);
int pic_particle_count(pic_ctx *ctx, int species, int64_t *n);
/* FP32 form only: the raw stored representation: cell = (k<<20)|(l<<10)|m
 * in guard-inclusive indices (-1 for a dead slot), offsets in [0,1) so that
 * x = (k + ox) dx, and velocities. */
int pic_download_particles_raw32(pic_ctx *ctx, int species, int64_t cap,
                                 int64_t *n, int32_t *cell, float *ox,
                                 float *oy, float *oz, float *vx, float *vy,
                                 float *vz);

/* ---- step routines (one per reference routine, in step order) ---- */

/* init_particle_flat (:412-625): ppc particles per interior cell, cells in
 * (k,l,m) order, uniform in the cell, drifting Maxwellian per species. */
int pic_load_particles(pic_ctx *ctx);
/* renew_grid (:2533-2578): zero J and n. */
int pic_renew_grid(pic_ctx *ctx);
/* renew_b4_2 (:2580-2611): half-step B with the (2,4) curl of E. */
int pic_renew_b_half(pic_ctx *ctx);
/* sr_b (:1001-1193) / sr_e (:1195-1387) on one rank: periodic guard fill
 * along x then y; z guards are left untouched. */
int pic_halo_b(pic_ctx *ctx);
int pic_halo_e(pic_ctx *ctx);
/* Hot loop A (:731-812): ballistic drift or cal_track4 (:2804-2870) with
 * shape_func_v2 gather (:3017-3326), then J deposit (:3397-3553). */
int pic_push_deposit_j(pic_ctx *ctx, int species);
/* Hot loop B (:818-867): x += v dt/2 and density deposit (:3333-3395). */
int pic_drift_deposit_n(pic_ctx *ctx, int species);
/* sr_current2 (:1805-2039) on one rank: additive periodic guard fold. */
int pic_fold_current(pic_ctx *ctx);
/* renew_e4_2 (:2613-2652). */
int pic_renew_e(pic_ctx *ctx);
/* sr_particle (:2041-2531) on one rank: drop particles outside the z range,
 * wrap x then y periodically (coordinate shift by the box length). */
int pic_particle_boundary(pic_ctx *ctx);
/* add_ion_flat/add_electron_flat (:5993-6817) for both species: drop the
 * particles in the z boundary slabs and refill 2-cell slabs. */
int pic_inflow(pic_ctx *ctx);
/* sorting (:7475-7590): stable per-cell sort, key (k,l,m) with m fastest,
 * particles outside the interior are dropped. */
int pic_sort(pic_ctx *ctx, int species);
/* Diagnostics (thrust_f :4940-5122 + Gauss residual :8345-8348).  Syncs. */
int pic_diagnose(pic_ctx *ctx, pic_diag *out);
/* No counterpart in the reference: pic_diagnose for callers that cannot
 * allocate (code under HIP graph capture, loops that forbid hipMalloc).
 * Writes the same record to dev_out, device memory of pic_sizeof_diag()
 * bytes owned by the caller, asynchronously on the context stream; calls no
 * allocator and does not synchronise.  Its workgroup partials and counters
 * live in a static device array of 95,234 doubles (744 KiB) per GPU, held
 * until the process ends; the layout is fixed, so every context fits.  All
 * calls on one GPU share that array: calls that overlap on one GPU (several
 * contexts' streams at once) race. */
int pic_diagnose_device(pic_ctx *ctx, pic_diag *dev_out);

/* fdtd() step loop (:687-949): nsteps full steps with the configured sort
 * and inflow cadences, fusing loops A and B and the particle boundary into
 * one particle pass per species.  FP32 form: a step's tile sort copies no
 * particles, the species' pass of that step reads them through the sort's
 * permutation; on a step with both a sort and an inflow the inflow runs
 * first (the same particles as the reference's order, in other slots). */
int pic_step(pic_ctx *ctx, int nsteps);
/* thrust_fave (:926, :5124-5285), every step of pic_step: the Lorentz force
 * q (E + v x B) n dx^3 on the particles of the central box (0.2 N dx <
 * x, y, z < 0.8 N dx) at the end of each step, summed over the steps since
 * the last reset.  The force of step c is added during the particle pass
 * of step c + 1 (same positions, velocities and E; the end-of-step B is
 * kept before the first B half-step), so the summation costs no extra
 * particle stream.  pic_thrust_ave_enable(on) resets the sum.
 * pic_thrust_ave_get fills the reference's F[27] (F[0..2] total, F[3..8]
 * ion E and v x B parts, F[9..14] electron parts, F[15..26] zero: every
 * particle has flag 1) with the sum over *nsteps steps (the reference
 * writes Fave = sum / 500 every 500 steps) and syncs; reset zeroes the sum.
 * Routines called one by one add nothing: the sum covers pic_step only. */
int pic_thrust_ave_enable(pic_ctx *ctx, int on);
int pic_thrust_ave_get(pic_ctx *ctx, double F[27], int64_t *nsteps, int reset);
/* The reference's pre-loop inflow (:673-674) followed by nothing else;
 * call once after pic_load_particles and before the first pic_step. */
int pic_begin(pic_ctx *ctx);
/* renew_phi3 (:8315-8665), sr_phi (:1389-1573, with the E correction) and
 * sr_e: the Dirichlet Poisson correction of E (source q(ni-ne)/eps0 - div E
 * with the (2,4) divergence, sine transforms along z, y, x, division by the
 * 7-point Laplacian symbol, inverse transforms; E += (2,4) gradient of phi).
 * PIC_E_STATE when poisson_every is 0 (no buffers). */
int pic_poisson(pic_ctx *ctx);

/* Species sorts done so far by each method: the FP32 form sorts with the
 * tile sort (particles still tile-sorted since the last sort, each tile's
 * movers within its list) and otherwise, as the FP64 form always does, with
 * a radix sort of (key, slot) pairs.  Both give the stable key order. */
int pic_sort_stats(const pic_ctx *ctx, int64_t *tile, int64_t *radix);

/* Step counter (c) and cadence counters (c_add, c_thrust).  pic_step keeps
 * them; a host loop that calls the routines itself keeps its own (fdtd's
 * locals, :630-686) and sets them before pic_checkpoint_save. */
int pic_get_counters(const pic_ctx *ctx, int64_t *step, int32_t *c_add,
                     int32_t *c_sort);
int pic_set_counters(pic_ctx *ctx, int64_t step, int32_t c_add, int32_t c_sort);

/* ---- kernel timing (HIP events on the context stream) ---- */
#define PIC_T_PUSH 0     /* fused particle pass */
#define PIC_T_FIELD_B 1  /* B half-step */
#define PIC_T_FIELD_E 2  /* E step */
#define PIC_T_SORT 3
#define PIC_T_INFLOW 4
#define PIC_T_OTHER 5    /* halos, fold, zeroing */
#define PIC_T_PUSH_SORT 6 /* fused particle pass applying pic_step's sort
                             permutation (FP32 form) */
#define PIC_T_COUNT 7
/* on: 0 off, 1 every class, 2 the particle-pass classes only (PIC_T_PUSH,
 * PIC_T_PUSH_SORT); every event record adds a packet to the stream, so mode
 * 2 leaves the step's other launches back to back. */
int pic_timing_enable(pic_ctx *ctx, int on);
/* Total milliseconds and launch count accumulated for a timing class. */
int pic_timing_get(pic_ctx *ctx, int which, double *ms, int64_t *launches);
int pic_timing_reset(pic_ctx *ctx);

/*
 * Adjoint of the step (no counterpart in the reference).  Gradient of the
 * loss L = (1/N) sum over the N steps of the z (flow-direction) flux of the
 * Maxwell stress tensor, eps0 (E_z E.n - n_z |E|^2/2) + (B_z B.n - n_z
 * |B|^2/2)/mu0 with the total B, through a sphere of `radius` cells around
 * the dipole (nquad Fibonacci points, equal weights), with respect to
 * p = (Mx, My, Mz, V): the dipole moment vector (B0 = M-field + IMF) and
 * the inflow drift speed along z.
 *
 * pic_adj_forward sets B0 from M and V, loads the particles and runs nsteps
 * steps of the context's form (the cell sort at its cadence; each inflow
 * appended to the store, whose capacity must hold it), checkpointing the
 * state every ckpt_every steps; it returns L.  pic_adj_backward recomputes
 * each segment from its checkpoint, recording each sort's permutation, and
 * reverses it in FP64 (transposed FDTD, guard copies, fold, gather, Boris
 * step and deposit, the particle adjoints carried back across each sort),
 * returning dloss * dL/dp.  The context is left at the end of the last
 * recomputed segment.
 */
int pic_adj_forward(pic_ctx *ctx, const double params[4], int nsteps,
                    int ckpt_every, double radius, int nquad, double *loss);
int pic_adj_backward(pic_ctx *ctx, double dloss, double grad[4]);

/* ---- field output ----
 *
 * Cell-centred text dumps of the fields, one line per cell, "%.3E" columns
 * separated by single spaces, cells in (k,l,m) order with m fastest:
 * - PIC_OUTPUT_COLUMNS, the reference's output() (:4834-4883): the
 *   interior cells, 23 columns: x, y, z of the cell corner relative to the
 *   box centre; ni, ne, ni - ne; E averaged over the 2 nodes along its
 *   axis; total B averaged over the 4 nodes of its staggered face; the six
 *   J node values; phi (PIC_PHI: 0 unless the Poisson correction ran in
 *   this step); |B|; B - B0 averaged;
 * - PIC_OUTPUT_TECPLOT, output_tecplot() (:4885-4938): a VARIABLES and a
 *   ZONE header line (step in T and SOLUTIONTIME), cells k, l up to N+2
 *   inclusive, 24 columns (J averaged along its axis and |J_i - J_e| added),
 *   an empty line after each (k,l) column.
 * Either numeric form; the step only enters the Tecplot header. */
#define PIC_OUTPUT_COLUMNS 0
#define PIC_OUTPUT_TECPLOT 1
int pic_output_fields(pic_ctx *ctx, const char *path, int format, int64_t step);

/* ---- checkpoint and resume ----
 *
 * Replaces the reference's suspend_job / continue_job (:7629-7911), which
 * dump the raw Grid records and Particle structs (pointers included) and
 * re-seed the random streams on resume.  A checkpoint holds the complete
 * state of a context between steps in its own numeric form: the 17 field
 * planes as stored (FP32 form: B - B0 in the B planes, plus B0 in double),
 * the particle slots of both species (tombstones included, in slot order),
 * the tile ranges of the last sort, the step counters and the Philox event
 * counter of the inflow, so a resumed run continues exactly as the original
 * (bit-identically in the FP32 deterministic mode; to atomic summation
 * order otherwise).  FP64-form checkpoints are also read and written by the
 * CPU oracle.
 *
 * File layout (little-endian): pic_checkpoint_header (256 bytes), then the
 * payload: the pic_config (config_bytes), 17 planes of (nx+4)(ny+4)(nz+4)
 * values in the order of enum pic_field (double, or float in the FP32
 * form), FP32 form only: 3 double planes of B0; then per species either six
 * double arrays x, y, z, vx, vy, vz (FP64) or an int32 cell array and six
 * float arrays ox, oy, oz, vx, vy, vz (FP32) of n[s] values; then per
 * species ntiles + 1 uint32 tile starts (none if ntiles is 0).  checksum is
 * a 64-bit hash of the payload: four interleaved FNV-1a-style lanes
 * h_j = (h_j ^ w) * 0x100000001b3 over the little-endian 64-bit words w of
 * the payload (word i to lane i mod 4, the last partial word zero-padded),
 * lanes seeded 0xcbf29ce484222325 + j and combined as
 * h = ((h_0 * 31 + h_1) * 31 + h_2) * 31 + h_3.
 *
 * pic_checkpoint_load needs a context created with the same config (the
 * capacity and the deposit mode may differ) holding at least n[s] slots.
 * If the file's tile geometry differs from the context's, the particles are
 * treated as unsorted until the next sort.  On PIC_E_FORMAT from a checksum
 * mismatch the state has been overwritten: load again or recreate. */
typedef struct pic_checkpoint_header {
  char magic[8];           /* "PLASCKPT" */
  uint32_t version;        /* 1 */
  uint32_t header_bytes;   /* 256 */
  uint32_t config_bytes;   /* sizeof(pic_config) of the writer */
  int32_t precision;       /* PIC_FP64 or PIC_FP32 */
  int32_t nx, ny, nz;
  int32_t tile[3];         /* tile geometry of the tile starts */
  int64_t step;            /* steps done (reference c) */
  int32_t c_add, c_sort;   /* inflow and sort cadence counters */
  uint32_t inflow_event;   /* Philox event id of the next inflow */
  uint32_t ntiles;         /* 0: no tile starts stored */
  int64_t n[2];            /* stored slots per species */
  int64_t sorted_end[2];   /* tile-sorted prefix per species */
  uint64_t payload_bytes;
  uint64_t checksum;
  int32_t drift_halves[2]; /* FP32 form: half-step drifts since the tile sort
                              per species (the tile windows' drift shift);
                              0 from writers without tile windows */
  uint8_t reserved[128];
} pic_checkpoint_header;

int pic_checkpoint_save(pic_ctx *ctx, const char *path);
int pic_checkpoint_load(pic_ctx *ctx, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* PLASMASIM_H */
